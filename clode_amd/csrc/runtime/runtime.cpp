// Logging, registries, device buffers and the hipRTC program cache.
//
// Program assembly follows the reference's scheme (clode/cpp/CLODE.cpp:109-152,
// CLODEfeatures.cpp:34-39): compile-time defines for precision, stepper,
// observer and problem sizes, then the user's getRHS source, then the
// drivers.  The reference builds OpenCL C with clBuildProgram
// (clode/cpp/OpenCLResource.cpp:260-296); here hipRTC builds a gfx950 code
// object which is cached by content hash in memory and on disk.
#include "clode_runtime.h"

#include <hip/hiprtc.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <cstdlib>
#include <sstream>
#include <dlfcn.h>
#include <hip/hip_version.h>
#include <sys/stat.h>
#include <unistd.h>

namespace clode {

// embedded device headers: kDeviceSources[] = {name, text}
#include "device_sources.inc"

// ------------------------------------------------------------------ logging
static int g_log_level = LOG_WARN;
static std::mutex g_log_mu;

void set_log_level(int level) { g_log_level = level; }
int get_log_level() { return g_log_level; }

void log_msg(int level, const std::string &msg)
{
    if (level < g_log_level || level >= LOG_OFF)
        return;
    static const char *names[] = {"trace", "debug", "info", "warning", "error", "critical"};
    std::lock_guard<std::mutex> lk(g_log_mu);
    std::fprintf(stderr, "[clode] [%s] %s\n", names[level], msg.c_str());
}

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) {
        std::string m = std::string(what) + ": " + hipGetErrorString(e);
        log_msg(LOG_ERR, m);
        throw RuntimeError(m);
    }
}

// --------------------------------------------------------------- registries
const std::map<std::string, int> &stepper_map()
{
    static const std::map<std::string, int> m = {
        {"euler", 1}, {"heun", 2}, {"rk4", 3}, {"bs23", 4}, {"dopri5", 5}, {"seuler", 6},
        // adaptive_he12.clh: compiled by the reference under HEUN_EULER but
        // missing from its name map (steppers.cl:26-43)
        {"he12", 7}};
    return m;
}

const std::map<std::string, int> &observer_map()
{
    static const std::map<std::string, int> m = {
        {"thresh2", CLODE_OBS_THRESH2}, {"localmax", CLODE_OBS_LOCALMAX}, {"nhood2", CLODE_OBS_NHOOD2},
        {"basic", CLODE_OBS_BASIC},     {"basicall", CLODE_OBS_BASICALL}, {"nhood1", CLODE_OBS_NHOOD1}};
    return m;
}

// feature names, in the order the device writes them
// (observer_threshold_2.clh:15-53, observer_local_maximum.clh:19-45,
//  observer_neighborhood_2.clh:14-41, observer_basic.clh:6-14,
//  observer_basic_allVar.clh:6-27, observer_neighborhood_1.clh:6-37)
std::vector<std::string> observer_feature_names(int obs, const ProblemInfo &pi, int nStore, unsigned int fVar)
{
    std::vector<std::string> f;
    auto var = [&](int j) { return j < (int)pi.varNames.size() ? pi.varNames[j] : "x" + std::to_string(j); };
    auto aux = [&](int j) { return j < (int)pi.auxNames.size() ? pi.auxNames[j] : "aux" + std::to_string(j); };
    if (obs == CLODE_OBS_THRESH2) {
        for (const char *s : {"period", "peaks", "upDuration", "downDuration", "duty", "activeDip"})
            for (const char *p : {"max ", "min ", "mean "})
                f.push_back(std::string(p) + s);
        for (int j = 0; j < pi.nVar; ++j) {
            f.push_back("max " + var(j));
            f.push_back("min " + var(j));
            f.push_back("mean " + var(j));
            f.push_back("max d" + var(j) + "/dt");
            f.push_back("min d" + var(j) + "/dt");
        }
        for (int j = 0; j < pi.nAux; ++j) {
            f.push_back("max " + aux(j));
            f.push_back("min " + aux(j));
            f.push_back("mean " + aux(j));
        }
        for (int j = 0; j < nStore; ++j) {
            f.push_back("up event time " + std::to_string(j));
            f.push_back("down event time " + std::to_string(j));
        }
        for (const char *s : {"event count", "step count", "max dt", "min dt", "mean dt"})
            f.push_back(s);
    } else if (obs == CLODE_OBS_LOCALMAX) {
        for (const char *s : {"IMI", "amplitude"})
            for (const char *p : {"max ", "min ", "mean "})
                f.push_back(std::string(p) + s);
        for (int j = 0; j < pi.nVar; ++j) {
            f.push_back("max " + var(j));
            f.push_back("min " + var(j));
            f.push_back("mean " + var(j));
            f.push_back("max d" + var(j) + "/dt");
            f.push_back("min d" + var(j) + "/dt");
        }
        for (int j = 0; j < pi.nAux; ++j) {
            f.push_back("max " + aux(j));
            f.push_back("min " + aux(j));
            f.push_back("mean " + aux(j));
        }
        for (int j = 0; j < nStore; ++j) {
            f.push_back("localmax event time " + std::to_string(j));
            f.push_back("localmax event evar " + std::to_string(j));
            f.push_back("localmin event time " + std::to_string(j));
            f.push_back("localmin event evar " + std::to_string(j));
        }
        f.push_back("event count");
        f.push_back("step count");
    } else if (obs == CLODE_OBS_NHOOD2) {
        for (const char *s : {"period", "peaks"})
            for (const char *p : {"max ", "min ", "mean "})
                f.push_back(std::string(p) + s);
        for (int j = 0; j < pi.nVar; ++j) {
            f.push_back("max " + var(j));
            f.push_back("min " + var(j));
            f.push_back("mean " + var(j));
            f.push_back("range " + var(j));
            f.push_back("nhood center " + var(j));
            f.push_back("max d" + var(j) + "/dt");
            f.push_back("min d" + var(j) + "/dt");
        }
        for (int j = 0; j < pi.nAux; ++j) {
            f.push_back("max " + aux(j));
            f.push_back("min " + aux(j));
            f.push_back("mean " + aux(j));
        }
        for (int j = 0; j < nStore; ++j)
            f.push_back("nhood event time " + std::to_string(j));
        for (const char *s : {"event count", "step count", "max dt", "min dt", "mean dt"})
            f.push_back(s);
    } else if (obs == CLODE_OBS_BASIC) {
        // basic reports the feature variable only
        int k = (int)fVar;
        f.push_back("max " + var(k));
        f.push_back("min " + var(k));
        f.push_back("mean " + var(k));
        f.push_back("max d" + var(k) + "/dt");
        f.push_back("min d" + var(k) + "/dt");
        f.push_back("step count");
    } else if (obs == CLODE_OBS_BASICALL || obs == CLODE_OBS_NHOOD1) {
        if (obs == CLODE_OBS_NHOOD1)
            for (const char *s : {"period", "peaks"})
                for (const char *p : {"max ", "min ", "mean "})
                    f.push_back(std::string(p) + s);
        for (int j = 0; j < pi.nVar; ++j) {
            f.push_back("max " + var(j));
            f.push_back("min " + var(j));
            f.push_back("mean " + var(j));
            f.push_back("max d" + var(j) + "/dt");
            f.push_back("min d" + var(j) + "/dt");
        }
        for (int j = 0; j < pi.nAux; ++j) {
            f.push_back("max " + aux(j));
            f.push_back("min " + aux(j));
            f.push_back("mean " + aux(j));
        }
        if (obs == CLODE_OBS_NHOOD1)
            for (const char *s : {"period count", "step count", "max dt", "min dt", "mean dt"})
                f.push_back(s);
        else
            f.push_back("step count");
    }
    return f;
}

// ------------------------------------------------------------------- device
int device_count()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess)
        return 0;
    return n;
}

void set_device(int device)
{
    int n = device_count();
    if (n <= 0)
        throw RuntimeError("no HIP device is available");
    if (device < 0 || device >= n)
        throw InvalidArgument("device ordinal " + std::to_string(device) + " out of range (" +
                              std::to_string(n) + " devices)");
    hip_check(hipSetDevice(device), "hipSetDevice");
}

void DeviceBuffer::allocate(int dev, size_t nbytes)
{
    release();
    if (nbytes == 0)
        nbytes = 8;
    set_device(dev);
    hip_check(hipMalloc(&ptr, nbytes), "hipMalloc");
    bytes = nbytes;
    device = dev;
}

void DeviceBuffer::release()
{
    if (ptr) {
        hipSetDevice(device);
        hipFree(ptr);
    }
    ptr = nullptr;
    bytes = 0;
}

// ------------------------------------------------------------------ program
std::string compose_source(const ProgramConfig &cfg)
{
    std::ostringstream s;
    s << "// clode_amd hipRTC program\n";
    s << "#define CLODE_PREC " << cfg.prec << "\n";
    s << "#define CLODE_STEPPER " << cfg.stepper << "\n";
    if (cfg.observer > 0) {
        s << "#define CLODE_OBSERVER " << cfg.observer << "\n";
        s << "#define CLODE_EVAR " << cfg.eVar << "\n";
        s << "#define CLODE_FVAR " << cfg.fVar << "\n";
    }
    s << "#define N_VAR " << cfg.nVar << "\n";
    s << "#define N_PAR " << cfg.nPar << "\n";
    s << "#define N_AUX " << cfg.nAux << "\n";
    s << "#define N_WIENER " << cfg.nWiener << "\n";
    s << "#define N_STORE_EVENTS " << cfg.nStore << "\n";
    if (cfg.wavesPerEU > 0)
        s << "#define CLODE_KERNEL_ATTRS __attribute__((amdgpu_waves_per_eu(" << cfg.wavesPerEU << ")))\n";
    if (cfg.replay)
        s << "#define CLODE_REPLAY 1\n";
    s << "#include \"clode_prelude.h\"\n";
    s << "#pragma clang force_cuda_host_device begin\n";
    // basename only: the program text (and its cache key) must not depend on
    // where the model file lives
    std::string rname = cfg.rhs_name.empty() ? std::string("rhs.cl") : cfg.rhs_name;
    size_t slash = rname.find_last_of('/');
    if (slash != std::string::npos)
        rname = rname.substr(slash + 1);
    s << "#line 1 \"" << rname << "\"\n";
    s << cfg.rhs << "\n";
    s << "#pragma clang force_cuda_host_device end\n";
    s << "#include \"clode_rng.h\"\n";
    s << "#include \"clode_steppers.h\"\n";
    if (cfg.observer > 0)
        s << "#include \"clode_observers.h\"\n";
    s << "#include \"clode_kernels.h\"\n";
    s << "#include \"clode_queue.h\"\n";
    return s.str();
}

// One system of a batched program: its defines, right-hand side, steppers,
// observer and drivers inside namespace clode_sys_<k>.
static void compose_system(std::ostringstream &s, const ProgramConfig &cfg, size_t k)
{
    s << "namespace clode_sys_" << k << " {\n";
    s << "#define CLODE_STEPPER " << cfg.stepper << "\n";
    if (cfg.observer > 0) {
        s << "#define CLODE_OBSERVER " << cfg.observer << "\n";
        s << "#define CLODE_EVAR " << cfg.eVar << "\n";
        s << "#define CLODE_FVAR " << cfg.fVar << "\n";
    }
    s << "#define N_VAR " << cfg.nVar << "\n";
    s << "#define N_PAR " << cfg.nPar << "\n";
    s << "#define N_AUX " << cfg.nAux << "\n";
    s << "#define N_WIENER " << cfg.nWiener << "\n";
    s << "#define N_STORE_EVENTS " << cfg.nStore << "\n";
    s << "#pragma clang force_cuda_host_device begin\n";
    std::string rname = cfg.rhs_name.empty() ? std::string("rhs.cl") : cfg.rhs_name;
    size_t slash = rname.find_last_of('/');
    if (slash != std::string::npos)
        rname = rname.substr(slash + 1);
    s << "#line 1 \"" << rname << "\"\n";
    s << cfg.rhs << "\n";
    s << "#pragma clang force_cuda_host_device end\n";
    s << "#include \"clode_steppers.h\"\n";
    if (cfg.observer > 0)
        s << "#include \"clode_observers.h\"\n";
    s << "#include \"clode_kernels.h\"\n";
    s << "}  // namespace clode_sys_" << k << "\n";
    s << "#include \"clode_system_end.h\"\n";
}

// Batched program: every system compiled once in its own namespace, then the
// dispatching work-queue kernels.  Wave w starts on system w mod nsys and
// visits the others in turn, so each wave works on one system at a time (no
// divergence between systems inside a wave) and waves whose system's queue is
// empty move on to help with the others.
std::string compose_batch_source(const std::vector<ProgramConfig> &cfgs)
{
    if (cfgs.empty() || cfgs.size() > CLODE_BATCH_MAX)
        throw InvalidArgument("a batch holds 1 to " + std::to_string(CLODE_BATCH_MAX) + " systems");
    const int prec = cfgs[0].prec;
    const bool obs = cfgs[0].observer > 0;
    for (const auto &c : cfgs) {
        if (c.prec != prec)
            throw InvalidArgument("all systems of a batch must use the same precision");
        if ((c.observer > 0) != obs)
            throw InvalidArgument("all systems of a batch must be feature simulators, or none");
    }
    const size_t n = cfgs.size();
    std::ostringstream s;
    s << "// clode_amd hipRTC batch program\n";
    s << "#define CLODE_PREC " << prec << "\n";
    s << "#define CLODE_BATCH 1\n";
    s << "#include \"clode_prelude.h\"\n";
    s << "#include \"clode_kernel_args.h\"\n";
    s << "#include \"clode_layout.h\"\n";
    s << "#include \"clode_rng.h\"\n";
    for (size_t k = 0; k < n; ++k)
        compose_system(s, cfgs[k], k);
    s << "#include \"clode_queue.h\"\n";
    s << "#define CLODE_NSYS " << n << "u\n";
    s << "struct ClodeBatchArgs {\n    ClodeArgs sys[CLODE_NSYS];\n};\n";
    auto dispatcher = [&](const char *name, const char *lane) {
        s << "extern \"C\" __global__ void __launch_bounds__(CLODE_BLOCK_MAX) clode_batch_" << name
          << "(const ClodeBatchArgs b)\n{\n";
        s << "    const unsigned int wave = (blockIdx.x * blockDim.x + threadIdx.x) / 64u;\n";
        s << "    for (unsigned int k = 0; k < CLODE_NSYS; ++k) {\n";
        s << "        switch ((wave + k) % CLODE_NSYS) {\n";
        for (size_t j = 0; j < n; ++j)
            s << "        case " << j << ": clode_sys_" << j << "::run_queue<clode_sys_" << j << "::" << lane
              << ">(b.sys[" << j << "]); break;\n";
        s << "        }\n    }\n}\n";
    };
    dispatcher("transient", "TransientLane");
    if (obs) {
        dispatcher("initialize_observer", "InitObsLane");
        dispatcher("features", "FeaturesLane");
    }
    // x0 <- xf for every system, grid-stride
    s << "extern \"C\" __global__ void __launch_bounds__(256) clode_batch_shift(const ClodeBatchArgs b)\n{\n";
    s << "    const size_t stride = (size_t)gridDim.x * blockDim.x;\n";
    s << "    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;\n";
    const char *xt = prec == PREC_SINGLE ? "float" : "double";
    for (size_t j = 0; j < n; ++j) {
        s << "    for (size_t e = t; e < (size_t)" << cfgs[j].nVar << " * b.sys[" << j << "].nPts; e += stride)\n";
        s << "        ((" << xt << " *)b.sys[" << j << "].x0)[e] = ((const " << xt << " *)b.sys[" << j
          << "].xf)[e];\n";
    }
    s << "}\n";
    return s.str();
}

std::vector<std::string> compile_options(const ProgramConfig &cfg)
{
    std::vector<std::string> o = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=on",
                                  // packed FP32 (v_pk_*_f32) is disabled for all device code
                                  "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                                  "-Wno-unused-parameter", "-Wno-unused-variable"};
    if (cfg.prec == PREC_FP32C) {
        o.push_back("-fno-hip-fp32-correctly-rounded-divide-sqrt");
        // float denormals flush to zero (double ones are kept): the float
        // division and exponential then need no denormal scaling or mode
        // switches
        o.push_back("-fgpu-flush-denormals-to-zero");
        // every division may multiply by the reciprocal (the hardware one
        // in float, a Newton-refined one in double)
        o.push_back("-freciprocal-math");
        // math functions may use the hardware approximations (exp and log
        // from v_exp_f32 and v_log_f32 without the library's extra-precision
        // steps)
        o.push_back("-fapprox-func");
        // the hardware's non-IEEE mode (which clang allows only with NaNs
        // assumed absent): min and max then need no canonicalising of their
        // operands, which IEEE mode spends an instruction on for every
        // loop-carried value.  Results of finite members are unchanged (the
        // features and final states of the lactotroph and Lorenz benchmark
        // ensembles compare bit for bit); NaN members may differ.
        o.push_back("-fno-honor-nans");
        o.push_back("-mno-amdgpu-ieee");
    } else
        o.push_back("-fhip-fp32-correctly-rounded-divide-sqrt");
    // the replay kernels unroll the features pass's observer once per ring
    // slot; at the default threshold the observer's own per-variable loops
    // then stay rolled, and their variable indices send the whole lane to
    // scratch memory
    if (cfg.replay) {
        o.push_back("-mllvm");
        o.push_back("-unroll-threshold=1000");
    }
    // code-generation tuning: CLODE_HIPRTC_OPTIONS holds extra whitespace-
    // separated compiler options (part of the program cache key)
    if (const char *extra = std::getenv("CLODE_HIPRTC_OPTIONS")) {
        std::istringstream in(extra);
        std::string tok;
        while (in >> tok)
            o.push_back(tok);
    }
    return o;
}

static uint64_t fnv1a(const char *p, size_t n, uint64_t h = 1469598103934665603ull)
{
    for (size_t i = 0; i < n; ++i) {
        h ^= (unsigned char)p[i];
        h *= 1099511628211ull;
    }
    return h;
}

static uint64_t fnv1a(const std::string &s) { return fnv1a(s.data(), s.size()); }

static std::mutex g_prog_mu;
// never destroyed: modules stay loaded until the process exits, so no HIP call
// runs during static destruction
static std::map<uint64_t, std::shared_ptr<Program>> &g_programs = *new std::map<uint64_t, std::shared_ptr<Program>>();

static std::string cache_path(const std::string &dir, uint64_t key)
{
    char name[64];
    std::snprintf(name, sizeof(name), "clode_%016llx.co", (unsigned long long)key);
    std::string d = dir;
    if (!d.empty() && d.back() != '/')
        d += '/';
    return d + name;
}

// cache file: "CLODECO2", the program key, the code size and an FNV-1a
// checksum of the code (each 8 bytes), then the code object.  A file whose
// header or checksum does not match is ignored and rebuilt.
static bool read_cache(const std::string &path, uint64_t key, std::vector<char> &code)
{
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f)
        return false;
    const std::streamoff fileSize = f.tellg();
    f.seekg(0);
    char magic[8];
    uint64_t k = 0, n = 0, sum = 0;
    f.read(magic, 8);
    f.read((char *)&k, 8);
    f.read((char *)&n, 8);
    f.read((char *)&sum, 8);
    if (!f || std::memcmp(magic, "CLODECO2", 8) != 0 || k != key)
        return false;
    // the header's size must account for the whole file before anything is
    // allocated: a truncated file or a damaged size field is rebuilt
    if (n == 0 || fileSize < 32 || (uint64_t)fileSize != 32 + n) {
        log_msg(LOG_WARN, "ignoring damaged cached code object " + path);
        return false;
    }
    std::vector<char> body(n);
    f.read(body.data(), (std::streamsize)n);
    if (!f || fnv1a(body.data(), n) != sum) {
        log_msg(LOG_WARN, "ignoring damaged cached code object " + path);
        return false;
    }
    code.swap(body);
    return true;
}

static void write_cache(const std::string &dir, const std::string &path, uint64_t key, const std::vector<char> &code)
{
    if (dir.empty())
        return;
    mkdir(dir.c_str(), 0755);
    std::string tmp = path + ".tmp" + std::to_string(getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (!f)
            return;
        const uint64_t n = code.size(), sum = fnv1a(code.data(), code.size());
        f.write("CLODECO2", 8);
        f.write((const char *)&key, 8);
        f.write((const char *)&n, 8);
        f.write((const char *)&sum, 8);
        f.write(code.data(), (std::streamsize)n);
    }
    std::ifstream check(tmp, std::ios::binary | std::ios::ate);
    if (!check || (uint64_t)check.tellg() != 32 + (uint64_t)code.size()) {
        std::remove(tmp.c_str());
        return;
    }
    std::rename(tmp.c_str(), path.c_str());
}

// Identity of the toolchain that builds code objects, part of every program
// key: the HIP release the library was built against (major.minor.patch and
// git hash), the version the HIP runtime reports, and the path, size and
// modification time of the hipRTC library actually loaded.  hipRTC's own
// version query returns a fixed API version, so it cannot tell two compilers
// apart; a ROCm upgrade or another install changes at least one of these.
// CLODE_TOOLCHAIN_ID (environment) is appended verbatim, so a user can force
// every cached code object to be rebuilt.
std::string toolchain_identity()
{
    std::ostringstream s;
    s << "hip " << HIP_VERSION_MAJOR << "." << HIP_VERSION_MINOR << "." << HIP_VERSION_PATCH << " "
      << HIP_VERSION_GITHASH;
    int rv = 0;
    if (hipRuntimeGetVersion(&rv) == hipSuccess)
        s << " runtime " << rv;
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&hiprtcCompileProgram), &info) && info.dli_fname) {
        s << " hiprtc " << info.dli_fname;
        struct stat st;
        if (stat(info.dli_fname, &st) == 0)
            s << " " << (long long)st.st_size << " " << (long long)st.st_mtime;
    }
    if (const char *e = std::getenv("CLODE_TOOLCHAIN_ID"))
        s << " id " << e;
    return s.str();
}

static void rtc_check(hiprtcResult r, const char *what)
{
    if (r != HIPRTC_SUCCESS)
        throw RuntimeError(std::string(what) + ": " + hiprtcGetErrorString(r));
}

std::shared_ptr<Program> get_program(const ProgramConfig &cfg, const std::string &cache_dir)
{
    return get_program_source(compose_source(cfg), compile_options(cfg), cache_dir);
}

std::shared_ptr<Program> get_program_source(const std::string &source, const std::vector<std::string> &options,
                                            const std::string &cache_dir)
{
    auto prog = std::make_shared<Program>();
    prog->source = source;
    prog->options = options;
    std::string keytext = prog->source;
    for (const auto &o : prog->options)
        keytext += "\n" + o;
    for (const auto &h : kDeviceSources)
        keytext += h.text;
    // a different compiler must not reuse code objects cached by another one
    keytext += "\n" + toolchain_identity();
    const uint64_t key = fnv1a(keytext);

    std::lock_guard<std::mutex> lk(g_prog_mu);
    auto it = g_programs.find(key);
    if (it != g_programs.end())
        return it->second;

    const std::string path = cache_dir.empty() ? std::string() : cache_path(cache_dir, key);
    if (!path.empty() && read_cache(path, key, prog->code)) {
        log_msg(LOG_DEBUG, "loaded cached code object " + path);
    } else {
        std::vector<const char *> hdrs, names;
        for (const auto &h : kDeviceSources) {
            hdrs.push_back(h.text);
            names.push_back(h.name);
        }
        hiprtcProgram p;
        rtc_check(hiprtcCreateProgram(&p, prog->source.c_str(), "clode_program.hip", (int)hdrs.size(),
                                      hdrs.data(), names.data()),
                  "hiprtcCreateProgram");
        std::vector<const char *> opts;
        for (const auto &o : prog->options)
            opts.push_back(o.c_str());
        hiprtcResult r = hiprtcCompileProgram(p, (int)opts.size(), opts.data());
        size_t logsize = 0;
        hiprtcGetProgramLogSize(p, &logsize);
        std::string buildlog(logsize, '\0');
        if (logsize > 1)
            hiprtcGetProgramLog(p, &buildlog[0]);
        if (r != HIPRTC_SUCCESS) {
            hiprtcDestroyProgram(&p);
            log_msg(LOG_ERR, "hipRTC build failed:\n" + buildlog);
            throw RuntimeError("hipRTC build failed: " + buildlog);
        }
        if (logsize > 1)
            log_msg(LOG_DEBUG, "hipRTC build log:\n" + buildlog);
        size_t n = 0;
        rtc_check(hiprtcGetCodeSize(p, &n), "hiprtcGetCodeSize");
        prog->code.resize(n);
        rtc_check(hiprtcGetCode(p, prog->code.data()), "hiprtcGetCode");
        hiprtcDestroyProgram(&p);
        if (!path.empty())
            write_cache(cache_dir, path, key, prog->code);
        log_msg(LOG_INFO, "compiled program (" + std::to_string(n) + " bytes)");
    }
    g_programs[key] = prog;
    return prog;
}

hipFunction_t Program::function(int device, const char *name)
{
    std::lock_guard<std::mutex> lk(mu_);
    set_device(device);
    hipModule_t mod;
    auto it = modules_.find(device);
    if (it == modules_.end()) {
        hip_check(hipModuleLoadData(&mod, code.data()), "hipModuleLoadData");
        modules_[device] = mod;
    } else {
        mod = it->second;
    }
    hipFunction_t fn;
    hip_check(hipModuleGetFunction(&fn, mod, name), (std::string("hipModuleGetFunction ") + name).c_str());
    return fn;
}

Program::~Program()
{
    for (auto &kv : modules_) {
        hipSetDevice(kv.first);
        hipModuleUnload(kv.second);
    }
}

}  // namespace clode
