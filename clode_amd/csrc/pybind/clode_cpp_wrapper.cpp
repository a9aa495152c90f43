// pybind11 module ``clode_cpp_wrapper``: the Python face of the C ABI.
//
// Mirrors the reference module of the same name (clode/cpp/CLODEpython.cpp:35-381):
// class names, method names, argument names and defaults are the same, so the
// reference's Python package logic can drive it.  Every call goes through
// include/clode_hip.h.  CLODE_ERR_INVALID raises ValueError and
// CLODE_ERR_RUNTIME raises RuntimeError, as the reference's
// std::invalid_argument / cl::Error do through pybind11.
//
// Device selection: the reference's OpenCLResource picks an OpenCL platform
// and device; here it picks a HIP device ordinal (platform ids are accepted and
// ignored, device ids are HIP ordinals).  Precision: the reference's bool
// cl_single_precision is accepted, and additionally the strings "single",
// "double" and "fp32c" (float state with Kahan-compensated updates).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "clode_hip.h"

namespace py = pybind11;

// 1-D float64 arrays: numpy arrays are used in place (lists and other
// sequences are converted), results are returned as owning numpy arrays
typedef py::array_t<double, py::array::c_style | py::array::forcecast> DArray;

namespace {

void check(int rc)
{
    if (rc == CLODE_OK)
        return;
    std::string msg = clode_last_error();
    if (rc == CLODE_ERR_INVALID)
        throw py::value_error(msg);
    throw std::runtime_error(msg);
}

std::string join(const std::vector<std::string> &v)
{
    std::string s;
    for (size_t i = 0; i < v.size(); ++i) {
        if (i)
            s += '\n';
        s += v[i];
    }
    return s;
}

std::vector<std::string> split(const std::string &s)
{
    std::vector<std::string> out;
    if (s.empty())
        return out;
    std::string cur;
    for (char c : s) {
        if (c == '\n') {
            out.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    out.push_back(cur);
    return out;
}

template <class Fn>
std::string get_string(Fn fn)
{
    size_t needed = 0;
    check(fn(nullptr, 0, &needed));
    std::string s(needed, '\0');
    check(fn(&s[0], needed, &needed));
    s.resize(needed ? needed - 1 : 0);
    return s;
}

// --------------------------------------------------------------- logging
enum LogLevelPy { trace = 0, debug, info, warn, err, critical, off };

struct Logger {
    void set_log_level(LogLevelPy l) { clode_set_log_level((int)l); }
    LogLevelPy get_log_level()
    {
        int l = 3;
        clode_get_log_level(&l);
        return (LogLevelPy)l;
    }
    void set_log_pattern(const std::string &) {}
};

Logger &get_logger()
{
    static Logger lg;
    return lg;
}

// ---------------------------------------------------------- data structs
struct ProblemInfoPy {
    std::string src_file;
    int num_var = 0, num_par = 0, num_aux = 0, num_noise = 0;
    std::vector<std::string> varNames, parNames, auxNames;
    ProblemInfoPy() = default;
    ProblemInfoPy(const std::string &f, const std::vector<std::string> &v, const std::vector<std::string> &p,
                  const std::vector<std::string> &a, int noise)
        : src_file(f), num_var((int)v.size()), num_par((int)p.size()), num_aux((int)a.size()), num_noise(noise),
          varNames(v), parNames(p), auxNames(a)
    {
    }
};

struct PiHolder {
    std::string vars, pars, aux;
    clode_problem_info pi;
    explicit PiHolder(const ProblemInfoPy &p)
        : vars(join(p.varNames)), pars(join(p.parNames)), aux(join(p.auxNames))
    {
        pi.src_file = p.src_file.c_str();
        pi.num_var = p.num_var;
        pi.num_par = p.num_par;
        pi.num_aux = p.num_aux;
        pi.num_noise = p.num_noise;
        pi.var_names = vars.c_str();
        pi.par_names = pars.c_str();
        pi.aux_names = aux.c_str();
    }
};

clode_solver_params default_sp()
{
    return clode_solver_params{0.1, 0.5, 1e-6, 1e-3, 1000000, 1000000, 1};
}

clode_observer_params default_op()
{
    return clode_observer_params{0, 0, 100, 0, 0., 0., 0.05, 0.2, 0.2, 0., 0., 0.};
}

// ------------------------------------------------------- device selector
enum CLDeviceTypePy {
    DEVICE_TYPE_ALL = 0,
    DEVICE_TYPE_CPU,
    DEVICE_TYPE_GPU,
    DEVICE_TYPE_ACCELERATOR,
    DEVICE_TYPE_DEFAULT,
    DEVICE_TYPE_CUSTOM
};
enum CLVendorPy { VENDOR_ANY = 0, VENDOR_NVIDIA, VENDOR_AMD, VENDOR_INTEL };

struct DeviceInfoPy {
    std::string name, vendor, version, devTypeStr, extensions;
    int devType = (int)DEVICE_TYPE_GPU;
    unsigned int computeUnits = 0, maxClock = 0;
    size_t maxWorkGroupSize = 0;
    unsigned long long deviceMemSize = 0, maxMemAllocSize = 0;
    bool doubleSupport = true, deviceAvailable = true;
};

struct PlatformInfoPy {
    std::string name = "HIP", vendor = "AMD", version;
    std::vector<DeviceInfoPy> device_info;
    unsigned int nDevices = 0;
};

DeviceInfoPy device_info(int d)
{
    clode_device_info di;
    check(clode_device_query(d, &di));
    DeviceInfoPy o;
    o.name = di.name;
    o.vendor = "Advanced Micro Devices, Inc.";
    o.version = std::string("HIP ") + di.arch;
    o.devTypeStr = "GPU";
    o.computeUnits = di.compute_units;
    o.maxClock = di.max_clock_khz / 1000;
    o.maxWorkGroupSize = di.max_work_group_size;
    o.deviceMemSize = di.device_memory_size;
    o.maxMemAllocSize = di.max_memory_alloc_size;
    o.doubleSupport = di.double_support != 0;
    o.extensions = di.arch;
    return o;
}

struct DeviceSelector {
    int device = 0;
    DeviceSelector() = default;
    explicit DeviceSelector(int d) : device(d) {}
    bool get_double_support(int = 0) { return true; }
    unsigned long long get_max_memory_alloc_size(int = 0)
    {
        int n = 0;
        clode_device_count(&n);
        if (n <= device)
            return 0;
        return device_info(device).maxMemAllocSize;
    }
    std::string get_device_cl_version(int = 0)
    {
        int n = 0;
        clode_device_count(&n);
        if (n <= device)
            return "";
        return device_info(device).version;
    }
    // OpenCLResource::print (OpenCLResource.cpp:313-319): info-level log
    // lines, so nothing is printed above LogLevel.info
    void print_devices()
    {
        int level = 0;
        clode_get_log_level(&level);
        if (level > 2)  // info
            return;
        int n = 0;
        clode_device_count(&n);
        py::print("Selected platform and device: ");
        py::print("Platform  --------------------");
        py::print("Name: HIP (AMD ROCm), selected through the OpenCLResource interface");
        for (int d = 0; d < n; ++d) {
            DeviceInfoPy i = device_info(d);
            py::print(d == device ? "* device" : "  device", d, ":", i.name, i.version, "CUs", i.computeUnits,
                      "memory", i.deviceMemSize);
        }
    }
};

std::vector<PlatformInfoPy> query_devices()
{
    PlatformInfoPy p;
    int n = 0;
    clode_device_count(&n);
    p.version = clode_version();
    for (int d = 0; d < n; ++d)
        p.device_info.push_back(device_info(d));
    p.nDevices = (unsigned int)n;
    return {p};
}

int parse_precision(const py::object &o)
{
    if (py::isinstance<py::bool_>(o))
        return o.cast<bool>() ? CLODE_PREC_SINGLE : CLODE_PREC_DOUBLE;
    if (py::isinstance<py::str>(o)) {
        std::string s = o.cast<std::string>();
        if (s == "single" || s == "fp32")
            return CLODE_PREC_SINGLE;
        if (s == "double" || s == "fp64")
            return CLODE_PREC_DOUBLE;
        if (s == "fp32c")
            return CLODE_PREC_FP32C;
        throw py::value_error("precision must be 'single', 'double' or 'fp32c'");
    }
    int v = o.cast<int>();
    if (v < 1 || v > 3)
        throw py::value_error("precision must be 1, 2 or 3");
    return v;
}

// --------------------------------------------------------------- simulators
class SimulatorBasePy {
public:
    SimulatorBasePy(const ProblemInfoPy &pi, const std::string &stepper, const py::object &single,
                    const DeviceSelector &dev, const std::string &root, int kind = CLODE_KIND_BASE,
                    const std::string &observer = "", const clode_observer_params *op = nullptr)
        : pi_(pi)
    {
        PiHolder h(pi);
        check(clode_create(&sim_, kind, &h.pi, stepper.c_str(), observer.empty() ? nullptr : observer.c_str(), op,
                           parse_precision(single), dev.device, root.empty() ? nullptr : root.c_str()));
    }
    virtual ~SimulatorBasePy() { clode_destroy(sim_); }

    void set_problem_info(const ProblemInfoPy &pi)
    {
        PiHolder h(pi);
        check(clode_set_problem_info(sim_, &h.pi));
        pi_ = pi;
    }
    void set_stepper(const std::string &s) { check(clode_set_stepper(sim_, s.c_str())); }
    void set_precision(const py::object &single) { check(clode_set_precision(sim_, parse_precision(single))); }
    void set_opencl(const DeviceSelector &d) { check(clode_set_device(sim_, d.device)); }
    void set_opencl_ids(unsigned int, unsigned int device) { check(clode_set_device(sim_, (int)device)); }
    void build_cl() { check(clode_build(sim_)); }
    void compile() { check(clode_compile(sim_)); }
    void set_problem_data(const DArray &x0, const DArray &pars)
    {
        check(clode_set_problem_data(sim_, x0.data(), (size_t)x0.size(), pars.data(), (size_t)pars.size()));
    }
    void set_tspan(const std::vector<double> &t) { check(clode_set_tspan(sim_, t.data(), t.size())); }
    void set_x0(const DArray &x) { check(clode_set_x0(sim_, x.data(), (size_t)x.size())); }
    void set_pars(const DArray &p) { check(clode_set_pars(sim_, p.data(), (size_t)p.size())); }
    void set_solver_params(const clode_solver_params &sp) { check(clode_set_solver_params(sim_, &sp)); }
    void seed_rng() { check(clode_seed_rng(sim_)); }
    void seed_rng_int(int seed) { check(clode_seed_rng_int(sim_, seed)); }
    void transient()
    {
        py::gil_scoped_release nogil;
        check(clode_transient(sim_));
    }
    void shift_tspan() { check(clode_shift_tspan(sim_)); }
    void shift_x0() { check(clode_shift_x0(sim_)); }
    ProblemInfoPy get_problem_info() const { return pi_; }
    std::vector<double> get_tspan()
    {
        std::vector<double> t(2);
        check(clode_get_tspan(sim_, t.data()));
        return t;
    }
    clode_solver_params get_solver_params()
    {
        clode_solver_params sp;
        check(clode_get_solver_params(sim_, &sp));
        return sp;
    }
    int n_pts()
    {
        int n = 0;
        check(clode_get_n_pts(sim_, &n));
        return n;
    }
    DArray get_pars()
    {
        DArray v((size_t)n_pts() * pi_.num_par);
        check(clode_get_pars(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    DArray get_x0()
    {
        DArray v((size_t)n_pts() * pi_.num_var);
        check(clode_get_x0(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    DArray get_xf()
    {
        DArray v((size_t)n_pts() * pi_.num_var);
        check(clode_get_xf(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    DArray get_dt()
    {
        DArray v((size_t)n_pts());
        check(clode_get_dt(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    DArray get_tf()
    {
        DArray v((size_t)n_pts());
        check(clode_get_tf(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    std::vector<std::string> get_available_steppers()
    {
        return split(get_string([&](char *b, size_t c, size_t *n) { return clode_get_available_steppers(sim_, b, c, n); }));
    }
    std::string get_program_string()
    {
        return get_string([&](char *b, size_t c, size_t *n) { return clode_get_program_string(sim_, b, c, n); });
    }
    void print_status() { check(clode_print_status(sim_)); }
    // extensions
    void set_launch(int mode, int block) { check(clode_set_launch(sim_, mode, block)); }
    void set_queue_slice(int slice) { check(clode_set_queue_slice(sim_, slice)); }
    clode_sim *handle() const { return sim_; }
    void set_member_tspan(const DArray &t)
    {
        check(clode_set_member_tspan(sim_, t.size() == 0 ? nullptr : t.data(), (size_t)t.size()));
    }
    void set_dt_refresh(bool on) { check(clode_set_dt_refresh(sim_, on ? 1 : 0)); }
    double last_kernel_ms()
    {
        double ms = 0;
        check(clode_last_kernel_ms(sim_, &ms));
        return ms;
    }
    int last_launch_variant()
    {
        int w = 0;
        check(clode_last_launch_variant(sim_, &w));
        return w;
    }
    int get_precision()
    {
        int p = 0;
        check(clode_get_precision(sim_, &p));
        return p;
    }
    void set_replay_budget(unsigned long long bytes) { check(clode_set_replay_budget(sim_, bytes)); }
    py::tuple get_replay_info()
    {
        unsigned int cap = 0;
        int replayed = 0;
        check(clode_get_replay_info(sim_, &cap, &replayed));
        return py::make_tuple(cap, replayed != 0);
    }
    void set_step_counter(bool on) { check(clode_set_step_counter(sim_, on ? 1 : 0)); }
    void reset_step_counts() { check(clode_reset_step_counts(sim_)); }
    std::vector<unsigned long long> get_step_counts()
    {
        std::vector<unsigned long long> v((size_t)n_pts());
        check(clode_get_step_counts(sim_, v.data(), v.size()));
        return v;
    }
    std::vector<unsigned long long> get_trial_counts()
    {
        std::vector<unsigned long long> v((size_t)n_pts());
        check(clode_get_trial_counts(sim_, v.data(), v.size()));
        return v;
    }
    py::array_t<unsigned long long> get_rng_state()
    {
        py::array_t<unsigned long long> v((py::ssize_t)(2 * (size_t)n_pts()));
        check(clode_get_rng_state(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }

protected:
    clode_sim *sim_ = nullptr;
    ProblemInfoPy pi_;
};

class FeatureSimulatorBasePy : public SimulatorBasePy {
public:
    FeatureSimulatorBasePy(const ProblemInfoPy &pi, const std::string &stepper, const std::string &observer,
                           const clode_observer_params &op, const py::object &single, const DeviceSelector &dev,
                           const std::string &root)
        : SimulatorBasePy(pi, stepper, single, dev, root, CLODE_KIND_FEATURES, observer, &op)
    {
    }
    void set_observer_params(const clode_observer_params &op) { check(clode_set_observer_params(sim_, &op)); }
    void set_observer(const std::string &o) { check(clode_set_observer(sim_, o.c_str())); }
    void initialize_observer()
    {
        py::gil_scoped_release nogil;
        check(clode_initialize_observer(sim_));
    }
    bool is_observer_initialized()
    {
        int v = 0;
        check(clode_is_observer_initialized(sim_, &v));
        return v != 0;
    }
    void transient_features(std::vector<double> tr, std::vector<double> fe)
    {
        if (tr.size() != 2 || fe.size() != 2)
            throw py::value_error("spans must have two elements");
        check(clode_transient_features(sim_, tr.data(), fe.data()));
    }
    void features_reinit(bool reinit)
    {
        py::gil_scoped_release nogil;
        check(clode_features(sim_, reinit ? 1 : 0));
    }
    void features()
    {
        py::gil_scoped_release nogil;
        check(clode_features(sim_, -1));
    }
    clode_observer_params get_observer_params()
    {
        clode_observer_params op;
        check(clode_get_observer_params(sim_, &op));
        return op;
    }
    std::string get_observer_name()
    {
        return get_string([&](char *b, size_t c, size_t *n) { return clode_get_observer_name(sim_, b, c, n); });
    }
    int get_n_features()
    {
        int n = 0;
        check(clode_get_n_features(sim_, &n));
        return n;
    }
    DArray get_f()
    {
        DArray v((size_t)get_n_features() * n_pts());
        check(clode_get_f(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    // F as a C-ordered (members, features) array, transposed on the device
    DArray get_f_members()
    {
        const py::ssize_t nf = get_n_features(), n = (py::ssize_t)n_pts();
        DArray v({n, nf});
        {
            py::gil_scoped_release nogil;
            check(clode_get_f_members(sim_, v.mutable_data(), (size_t)(n * nf)));
        }
        return v;
    }
    std::vector<std::string> get_feature_names()
    {
        return split(get_string([&](char *b, size_t c, size_t *n) { return clode_get_feature_names(sim_, b, c, n); }));
    }
    std::vector<std::string> get_available_observers()
    {
        return split(
            get_string([&](char *b, size_t c, size_t *n) { return clode_get_available_observers(sim_, b, c, n); }));
    }
};

class TrajectorySimulatorBasePy : public SimulatorBasePy {
public:
    TrajectorySimulatorBasePy(const ProblemInfoPy &pi, const std::string &stepper, const py::object &single,
                              const DeviceSelector &dev, const std::string &root)
        : SimulatorBasePy(pi, stepper, single, dev, root, CLODE_KIND_TRAJECTORY)
    {
    }
    void trajectory()
    {
        py::gil_scoped_release nogil;
        check(clode_trajectory(sim_));
    }
    void sizes(size_t &nt, size_t &nx, size_t &na) { check(clode_get_trajectory_sizes(sim_, &nt, &nx, &na)); }
    DArray get_t()
    {
        size_t nt, nx, na;
        sizes(nt, nx, na);
        DArray v(nt);
        check(clode_get_t(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    DArray get_x()
    {
        size_t nt, nx, na;
        sizes(nt, nx, na);
        DArray v(nx);
        check(clode_get_x(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    DArray get_dx()
    {
        size_t nt, nx, na;
        sizes(nt, nx, na);
        DArray v(nx);
        check(clode_get_dx(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    DArray get_aux()
    {
        size_t nt, nx, na;
        sizes(nt, nx, na);
        DArray v(na);
        check(clode_get_aux(sim_, v.mutable_data(), (size_t)v.size()));
        return v;
    }
    std::vector<int> get_n_stored()
    {
        std::vector<int> v((size_t)n_pts());
        check(clode_get_n_stored(sim_, v.data(), v.size()));
        return v;
    }
};

std::string vec_str(const std::vector<std::string> &v)
{
    std::string s = "[";
    for (size_t i = 0; i < v.size(); ++i)
        s += (i ? ", " : "") + v[i];
    return s + "]";
}

}  // namespace

// Several simulators in one work-queue launch (no reference counterpart); the
// batch keeps references to its simulators so they outlive it.
class BatchSimulatorBasePy {
public:
    explicit BatchSimulatorBasePy(const py::list &sims)
    {
        std::vector<clode_sim *> h;
        for (const auto &o : sims) {
            SimulatorBasePy &s = o.cast<SimulatorBasePy &>();
            h.push_back(s.handle());
            keep_.push_back(py::reinterpret_borrow<py::object>(o));
        }
        check(clode_batch_create(&b_, h.data(), h.size()));
    }
    ~BatchSimulatorBasePy() { clode_batch_destroy(b_); }
    BatchSimulatorBasePy(const BatchSimulatorBasePy &) = delete;
    BatchSimulatorBasePy &operator=(const BatchSimulatorBasePy &) = delete;

    void set_launch(int block, int slice) { check(clode_batch_set_launch(b_, block, slice)); }
    py::tuple get_launch()
    {
        int b = 0, s = 0;
        check(clode_batch_get_launch(b_, &b, &s));
        return py::make_tuple(b, s);
    }
    void transient()
    {
        py::gil_scoped_release nogil;
        check(clode_batch_transient(b_));
    }
    void initialize_observer()
    {
        py::gil_scoped_release nogil;
        check(clode_batch_initialize_observer(b_));
    }
    void features(int reinit)
    {
        py::gil_scoped_release nogil;
        check(clode_batch_features(b_, reinit));
    }
    void shift_x0()
    {
        py::gil_scoped_release nogil;
        check(clode_batch_shift_x0(b_));
    }
    void sweep(int rounds, bool graph)
    {
        py::gil_scoped_release nogil;
        check(clode_batch_sweep(b_, rounds, graph ? 1 : 0));
    }
    py::tuple autotune(int stride, unsigned int max_steps)
    {
        int b = 0, s = 0;
        {
            py::gil_scoped_release nogil;
            check(clode_batch_autotune(b_, stride, max_steps, &b, &s));
        }
        return py::make_tuple(b, s);
    }
    py::tuple autotune_table()
    {
        int n = 0, cached = 0;
        check(clode_batch_autotune_table(b_, nullptr, 0, &n, &cached));
        py::array_t<double> rows({(py::ssize_t)n, (py::ssize_t)5});
        if (n > 0)
            check(clode_batch_autotune_table(b_, rows.mutable_data(), n, &n, &cached));
        return py::make_tuple(rows, cached != 0);
    }
    double last_kernel_ms()
    {
        double ms = 0;
        check(clode_batch_last_ms(b_, &ms));
        return ms;
    }
    size_t size() const { return keep_.size(); }

private:
    clode_batch *b_ = nullptr;
    std::vector<py::object> keep_;
};

PYBIND11_MODULE(clode_cpp_wrapper, m)
{
    m.doc() = "clode_amd: MI355X ensemble ODE integrator (HIP runtime over the C ABI)";

    py::enum_<LogLevelPy>(m, "LogLevel")
        .value("trace", trace)
        .value("debug", debug)
        .value("info", info)
        .value("warn", warn)
        .value("err", err)
        .value("critical", critical)
        .value("off", off)
        .export_values();
    // bound under the reference's class name (clode/cpp/CLODEpython.cpp:87)
    py::class_<Logger>(m, "LoggerSingleton")
        .def("set_log_level", &Logger::set_log_level)
        .def("get_log_level", &Logger::get_log_level)
        .def("set_log_pattern", &Logger::set_log_pattern);
    m.def("get_logger", &get_logger, py::return_value_policy::reference);

    py::enum_<CLVendorPy>(m, "CLVendor")
        .value("VENDOR_ANY", VENDOR_ANY)
        .value("VENDOR_NVIDIA", VENDOR_NVIDIA)
        .value("VENDOR_AMD", VENDOR_AMD)
        .value("VENDOR_INTEL", VENDOR_INTEL)
        .export_values();
    py::enum_<CLDeviceTypePy>(m, "CLDeviceType")
        .value("DEVICE_TYPE_ALL", DEVICE_TYPE_ALL)
        .value("DEVICE_TYPE_CPU", DEVICE_TYPE_CPU)
        .value("DEVICE_TYPE_GPU", DEVICE_TYPE_GPU)
        .value("DEVICE_TYPE_ACCELERATOR", DEVICE_TYPE_ACCELERATOR)
        .value("DEVICE_TYPE_DEFAULT", DEVICE_TYPE_DEFAULT)
        .value("DEVICE_TYPE_CUSTOM", DEVICE_TYPE_CUSTOM)
        .export_values();

    py::class_<DeviceSelector>(m, "OpenCLResource")
        .def(py::init<>())
        .def(py::init([](CLDeviceTypePy) { return DeviceSelector(0); }))
        .def(py::init([](CLVendorPy) { return DeviceSelector(0); }))
        .def(py::init([](CLDeviceTypePy, CLVendorPy) { return DeviceSelector(0); }))
        .def(py::init([](unsigned int, unsigned int device) { return DeviceSelector((int)device); }))
        .def(py::init([](unsigned int, std::vector<unsigned int> devices) {
            return DeviceSelector(devices.empty() ? 0 : (int)devices[0]);
        }))
        .def_readonly("device", &DeviceSelector::device)
        .def("get_double_support", &DeviceSelector::get_double_support, py::arg("device_id") = 0)
        .def("get_max_memory_alloc_size", &DeviceSelector::get_max_memory_alloc_size, py::arg("device_id") = 0)
        .def("get_device_cl_version", &DeviceSelector::get_device_cl_version, py::arg("device_id") = 0)
        .def("print_devices", &DeviceSelector::print_devices);

    py::class_<DeviceInfoPy>(m, "DeviceInfo")
        .def_readwrite("name", &DeviceInfoPy::name)
        .def_readwrite("vendor", &DeviceInfoPy::vendor)
        .def_readwrite("version", &DeviceInfoPy::version)
        .def_readwrite("device_type", &DeviceInfoPy::devType)
        .def_readwrite("device_type_str", &DeviceInfoPy::devTypeStr)
        .def_readwrite("compute_units", &DeviceInfoPy::computeUnits)
        .def_readwrite("max_clock", &DeviceInfoPy::maxClock)
        .def_readwrite("max_work_group_size", &DeviceInfoPy::maxWorkGroupSize)
        .def_readwrite("device_memory_size", &DeviceInfoPy::deviceMemSize)
        .def_readwrite("max_memory_alloc_size", &DeviceInfoPy::maxMemAllocSize)
        .def_readwrite("extensions", &DeviceInfoPy::extensions)
        .def_readwrite("double_support", &DeviceInfoPy::doubleSupport)
        .def_readwrite("device_available", &DeviceInfoPy::deviceAvailable)
        .def("__repr__", [](const DeviceInfoPy &d) {
            return "<device_info(name=" + d.name + ", version=" + d.version +
                   ", compute_units=" + std::to_string(d.computeUnits) +
                   ", device_memory_size=" + std::to_string(d.deviceMemSize) + ")>";
        });
    py::class_<PlatformInfoPy>(m, "PlatformInfo")
        .def_readwrite("name", &PlatformInfoPy::name)
        .def_readwrite("vendor", &PlatformInfoPy::vendor)
        .def_readwrite("version", &PlatformInfoPy::version)
        .def_readwrite("device_info", &PlatformInfoPy::device_info)
        .def_readwrite("device_count", &PlatformInfoPy::nDevices)
        .def("__repr__", [](const PlatformInfoPy &p) {
            return "<platform_info(name=" + p.name + ", vendor=" + p.vendor + ", version=" + p.version +
                   ", device_count=" + std::to_string(p.nDevices) + ")>";
        });
    m.def("query_opencl", &query_devices, "Query HIP devices");
    m.def("_print_opencl", []() {
        for (auto &p : query_devices())
            for (auto &d : p.device_info)
                py::print(d.name, d.version, "CUs", d.computeUnits);
    });

    py::class_<ProblemInfoPy>(m, "ProblemInfo")
        .def(py::init<const std::string &, const std::vector<std::string> &, const std::vector<std::string> &,
                      const std::vector<std::string> &, int>(),
             py::arg("src_file"), py::arg("vars"), py::arg("pars"), py::arg("aux") = std::vector<std::string>(),
             py::arg("num_noise") = 1)
        .def(py::init<>())
        .def_readwrite("src_file", &ProblemInfoPy::src_file)
        .def_readwrite("num_var", &ProblemInfoPy::num_var)
        .def_readwrite("num_par", &ProblemInfoPy::num_par)
        .def_readwrite("num_aux", &ProblemInfoPy::num_aux)
        .def_readwrite("num_noise", &ProblemInfoPy::num_noise)
        .def_property(
            "vars", [](const ProblemInfoPy &p) { return p.varNames; },
            [](ProblemInfoPy &p, const std::vector<std::string> &v) {
                p.varNames = v;
                p.num_var = (int)v.size();
            })
        .def_property(
            "pars", [](const ProblemInfoPy &p) { return p.parNames; },
            [](ProblemInfoPy &p, const std::vector<std::string> &v) {
                p.parNames = v;
                p.num_par = (int)v.size();
            })
        .def_property(
            "aux", [](const ProblemInfoPy &p) { return p.auxNames; },
            [](ProblemInfoPy &p, const std::vector<std::string> &v) {
                p.auxNames = v;
                p.num_aux = (int)v.size();
            })
        .def("__repr__", [](const ProblemInfoPy &p) {
            return "<problem_info(src_file=" + p.src_file + ", vars=" + vec_str(p.varNames) +
                   ", pars=" + vec_str(p.parNames) + ", aux=" + vec_str(p.auxNames) +
                   ", num_noise=" + std::to_string(p.num_noise) + ")>";
        });

    py::class_<clode_solver_params>(m, "SolverParams")
        .def(py::init([](double dt, double dtmax, double abstol, double reltol, unsigned int max_steps,
                         unsigned int max_store, unsigned int nout) {
                 return clode_solver_params{dt, dtmax, abstol, reltol, max_steps, max_store, nout};
             }),
             py::arg("dt") = 0.1, py::arg("dtmax") = 0.5, py::arg("abstol") = 1e-6, py::arg("reltol") = 1e-3,
             py::arg("max_steps") = 1000000, py::arg("max_store") = 1000000, py::arg("nout") = 1)
        .def_readwrite("dt", &clode_solver_params::dt)
        .def_readwrite("dtmax", &clode_solver_params::dtmax)
        .def_readwrite("abstol", &clode_solver_params::abstol)
        .def_readwrite("reltol", &clode_solver_params::reltol)
        .def_readwrite("max_steps", &clode_solver_params::max_steps)
        .def_readwrite("max_store", &clode_solver_params::max_store)
        .def_readwrite("nout", &clode_solver_params::nout)
        .def("__repr__", [](const clode_solver_params &s) {
            return "<solver_params(dt=" + std::to_string(s.dt) + ", dtmax=" + std::to_string(s.dtmax) +
                   ", abstol=" + std::to_string(s.abstol) + ", reltol=" + std::to_string(s.reltol) +
                   ", max_steps=" + std::to_string(s.max_steps) + ", max_store=" + std::to_string(s.max_store) +
                   ", nout=" + std::to_string(s.nout) + ")>";
        });

    py::class_<clode_observer_params>(m, "ObserverParams")
        .def(py::init([](unsigned int e, unsigned int f, unsigned int mec, unsigned int met, double min_amp,
                         double min_imi, double nr, double xu, double xd, double dxu, double dxd, double eps) {
                 return clode_observer_params{e, f, mec, met, min_amp, min_imi, nr, xu, xd, dxu, dxd, eps};
             }),
             py::arg("e_var_ix") = 0, py::arg("f_var_ix") = 0, py::arg("max_event_count") = 100,
             py::arg("max_event_timestamps") = 0, py::arg("min_amp") = 0., py::arg("min_imi") = 0.,
             py::arg("nhood_radius") = 0.05, py::arg("x_up_threshold") = 0.2, py::arg("x_down_threshold") = 0.2,
             py::arg("dx_up_threshold") = 0., py::arg("dx_down_threshold") = 0., py::arg("eps_dx") = 0.)
        .def_readwrite("e_var_ix", &clode_observer_params::eVarIx)
        .def_readwrite("f_var_ix", &clode_observer_params::fVarIx)
        .def_readwrite("max_event_count", &clode_observer_params::maxEventCount)
        .def_readwrite("max_event_timestamps", &clode_observer_params::maxEventTimestamps)
        .def_readwrite("min_amp", &clode_observer_params::minXamp)
        .def_readwrite("min_imi", &clode_observer_params::minIMI)
        .def_readwrite("nhood_radius", &clode_observer_params::nHoodRadius)
        .def_readwrite("x_up_threshold", &clode_observer_params::xUpThresh)
        .def_readwrite("x_down_threshold", &clode_observer_params::xDownThresh)
        .def_readwrite("dx_up_threshold", &clode_observer_params::dxUpThresh)
        .def_readwrite("dx_down_threshold", &clode_observer_params::dxDownThresh)
        .def_readwrite("eps_dx", &clode_observer_params::eps_dx)
        .def("__repr__", [](const clode_observer_params &p) {
            return "<observer_params(e_var_ix=" + std::to_string(p.eVarIx) + ", f_var_ix=" + std::to_string(p.fVarIx) +
                   ", max_event_count=" + std::to_string(p.maxEventCount) +
                   ", max_event_timestamps=" + std::to_string(p.maxEventTimestamps) +
                   ", x_up_threshold=" + std::to_string(p.xUpThresh) +
                   ", x_down_threshold=" + std::to_string(p.xDownThresh) + ")>";
        });

    py::class_<SimulatorBasePy>(m, "SimulatorBase")
        .def(py::init<const ProblemInfoPy &, const std::string &, const py::object &, const DeviceSelector &,
                      const std::string &>(),
             py::arg("problem_info"), py::arg("stepper"), py::arg("cl_single_precision"), py::arg("opencl_resource"),
             py::arg("clode_root"))
        .def("set_problem_info", &SimulatorBasePy::set_problem_info)
        .def("set_stepper", &SimulatorBasePy::set_stepper)
        .def("set_precision", &SimulatorBasePy::set_precision)
        .def("set_opencl", &SimulatorBasePy::set_opencl)
        .def("set_opencl", &SimulatorBasePy::set_opencl_ids)
        .def("build_cl", &SimulatorBasePy::build_cl)
        .def("compile", &SimulatorBasePy::compile)
        .def("set_problem_data", &SimulatorBasePy::set_problem_data)
        .def("set_tspan", &SimulatorBasePy::set_tspan)
        .def("set_x0", &SimulatorBasePy::set_x0)
        .def("set_pars", &SimulatorBasePy::set_pars)
        .def("set_solver_params", &SimulatorBasePy::set_solver_params)
        .def("seed_rng", &SimulatorBasePy::seed_rng, "Seed RNG")
        .def("seed_rng", &SimulatorBasePy::seed_rng_int, "Seed RNG", py::arg("seed"))
        .def("transient", &SimulatorBasePy::transient)
        .def("shift_tspan", &SimulatorBasePy::shift_tspan)
        .def("shift_x0", &SimulatorBasePy::shift_x0)
        .def("get_problem_info", &SimulatorBasePy::get_problem_info)
        .def("get_tspan", &SimulatorBasePy::get_tspan)
        .def("get_solver_params", &SimulatorBasePy::get_solver_params)
        .def("get_pars", &SimulatorBasePy::get_pars)
        .def("get_x0", &SimulatorBasePy::get_x0)
        .def("get_xf", &SimulatorBasePy::get_xf)
        .def("get_dt", &SimulatorBasePy::get_dt)
        .def("get_tf", &SimulatorBasePy::get_tf)
        .def("get_available_steppers", &SimulatorBasePy::get_available_steppers)
        .def("get_program_string", &SimulatorBasePy::get_program_string)
        .def("print_status", &SimulatorBasePy::print_status)
        .def("get_n_pts", &SimulatorBasePy::n_pts)
        .def("get_precision", &SimulatorBasePy::get_precision)
        .def("set_launch", &SimulatorBasePy::set_launch, py::arg("mode"), py::arg("block_size") = 0)
        .def("set_queue_slice", &SimulatorBasePy::set_queue_slice)
        .def("set_member_tspan", &SimulatorBasePy::set_member_tspan)
        .def("set_dt_refresh", &SimulatorBasePy::set_dt_refresh)
        .def("last_kernel_ms", &SimulatorBasePy::last_kernel_ms)
        .def("last_launch_variant", &SimulatorBasePy::last_launch_variant)
        .def("set_replay_budget", &SimulatorBasePy::set_replay_budget)
        .def("get_replay_info", &SimulatorBasePy::get_replay_info)
        .def("set_step_counter", &SimulatorBasePy::set_step_counter)
        .def("reset_step_counts", &SimulatorBasePy::reset_step_counts)
        .def("get_step_counts", &SimulatorBasePy::get_step_counts)
        .def("get_trial_counts", &SimulatorBasePy::get_trial_counts)
        .def("get_rng_state", &SimulatorBasePy::get_rng_state);

    py::class_<FeatureSimulatorBasePy, SimulatorBasePy>(m, "FeatureSimulatorBase")
        .def(py::init<const ProblemInfoPy &, const std::string &, const std::string &, const clode_observer_params &,
                      const py::object &, const DeviceSelector &, const std::string &>())
        .def("build_cl", &FeatureSimulatorBasePy::build_cl)
        .def("set_observer_params", &FeatureSimulatorBasePy::set_observer_params)
        .def("set_observer", &FeatureSimulatorBasePy::set_observer)
        .def("initialize_observer", &FeatureSimulatorBasePy::initialize_observer)
        .def("is_observer_initialized", &FeatureSimulatorBasePy::is_observer_initialized)
        .def("features", &FeatureSimulatorBasePy::features_reinit)
        .def("features", &FeatureSimulatorBasePy::features)
        .def("transient_features", &FeatureSimulatorBasePy::transient_features, py::arg("transient_tspan"),
             py::arg("features_tspan"))
        .def("get_observer_params", &FeatureSimulatorBasePy::get_observer_params)
        .def("get_observer_name", &FeatureSimulatorBasePy::get_observer_name)
        .def("get_f", &FeatureSimulatorBasePy::get_f)
        .def("get_f_members", &FeatureSimulatorBasePy::get_f_members)
        .def("get_n_features", &FeatureSimulatorBasePy::get_n_features)
        .def("get_feature_names", &FeatureSimulatorBasePy::get_feature_names)
        .def("get_available_observers", &FeatureSimulatorBasePy::get_available_observers)
        .def("__repr__", [](FeatureSimulatorBasePy &c) {
            return "<CLODEfeatures (observer=" + c.get_observer_name() +
                   ", n_features=" + std::to_string(c.get_n_features()) + ")>";
        });

    py::class_<TrajectorySimulatorBasePy, SimulatorBasePy>(m, "TrajectorySimulatorBase")
        .def(py::init<const ProblemInfoPy &, const std::string &, const py::object &, const DeviceSelector &,
                      const std::string &>())
        .def("build_cl", &TrajectorySimulatorBasePy::build_cl)
        .def("trajectory", &TrajectorySimulatorBasePy::trajectory)
        .def("get_t", &TrajectorySimulatorBasePy::get_t)
        .def("get_x", &TrajectorySimulatorBasePy::get_x)
        .def("get_dx", &TrajectorySimulatorBasePy::get_dx)
        .def("get_aux", &TrajectorySimulatorBasePy::get_aux)
        .def("get_n_stored", &TrajectorySimulatorBasePy::get_n_stored);

    m.def(
        "compile_batch",
        [](const py::list &sims) {
            std::vector<clode_sim *> h;
            for (const auto &o : sims)
                h.push_back(o.cast<SimulatorBasePy &>().handle());
            check(clode_batch_compile(h.data(), h.size()));
        },
        "Compile the batched program for these simulators (no device needed)");

    py::class_<BatchSimulatorBasePy>(m, "BatchSimulatorBase")
        .def(py::init<const py::list &>(), py::arg("simulators"))
        .def("set_launch", &BatchSimulatorBasePy::set_launch, py::arg("block_size"), py::arg("slice"))
        .def("get_launch", &BatchSimulatorBasePy::get_launch)
        .def("transient", &BatchSimulatorBasePy::transient)
        .def("initialize_observer", &BatchSimulatorBasePy::initialize_observer)
        .def("features", &BatchSimulatorBasePy::features, py::arg("reinitialize") = -1)
        .def("shift_x0", &BatchSimulatorBasePy::shift_x0)
        .def("sweep", &BatchSimulatorBasePy::sweep, py::arg("rounds"), py::arg("graph") = true)
        .def("autotune", &BatchSimulatorBasePy::autotune, py::arg("stride") = 16, py::arg("max_steps") = 4096)
        .def("autotune_table", &BatchSimulatorBasePy::autotune_table)
        .def("last_kernel_ms", &BatchSimulatorBasePy::last_kernel_ms)
        .def("__len__", &BatchSimulatorBasePy::size);
}
