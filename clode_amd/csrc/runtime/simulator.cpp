// Host semantics of the reference's CLODE / CLODEfeatures / CLODEtrajectory
// classes on HIP.  Citations point at the reference member each method
// reproduces.
#include "simulator.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <random>
#include <sstream>

namespace clode {

static std::string read_file(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f)
        throw InvalidArgument("cannot read right-hand-side source file '" + path + "'");
    std::ostringstream s;
    s << f.rdbuf();
    return s.str();
}

// CLODE::CLODE (CLODE.cpp:13-37), CLODEfeatures::CLODEfeatures (CLODEfeatures.cpp:9-29),
// CLODEtrajectory::CLODEtrajectory (CLODEtrajectory.cpp:6-20)
Simulator::Simulator(int kind_, const ProblemInfo &pi, const std::string &stepper_, const std::string &observer_,
                     const ObserverParams &op_, int precision, int device_, const std::string &root_)
    : kind(kind_), prec(PREC_DOUBLE), device(device_), root(root_)
{
    if (precision < PREC_SINGLE || precision > PREC_FP32C)
        throw InvalidArgument("precision must be 1 (single), 2 (double) or 3 (fp32c)");
    prec = precision;
    setProblemInfo(pi);
    stepper = "rk4";
    setStepper(stepper_);
    if (stepper_map().find(stepper_) == stepper_map().end())
        throw InvalidArgument("unknown stepper: " + stepper_);
    if (kind == 1) {
        if (observer_map().find(observer_) == observer_map().end())
            throw InvalidArgument("unknown observer: " + observer_);
        observer = observer_;
        setObserverParams(op_);
    }
    log_msg(LOG_DEBUG, "constructed simulator");
}

Simulator::~Simulator()
{
    if (stream) {
        hipSetDevice(device);
        hipStreamSynchronize(stream);
        hipEventDestroy(ev0);
        hipEventDestroy(ev1);
        hipStreamDestroy(stream);
    }
}

// CLODE::setProblemInfo (CLODE.cpp:43-59)
void Simulator::setProblemInfo(const ProblemInfo &pi)
{
    ++inputGen;
    if (pi.nVar <= 0)
        throw InvalidArgument("the problem must have at least one state variable");
    if (pi.nPar < 0 || pi.nAux < 0 || pi.nWiener < 0)
        throw InvalidArgument("negative problem dimension");
    rhsSource = read_file(pi.src_file);
    prob = pi;
    if (kind == 1 && !observer.empty())
        updateObserverInfo();
}

// CLODE::setStepper (CLODE.cpp:61-76): unknown names leave the stepper unchanged
void Simulator::setStepper(const std::string &s)
{
    ++inputGen;
    if (stepper_map().find(s) != stepper_map().end())
        stepper = s;
    else
        log_msg(LOG_WARN, "Unknown stepper: " + s + ". Stepper method unchanged");
}

// CLODE::setPrecision (CLODE.cpp:78-86).  Device data of the old precision are
// released; the caller sets problem data again, as with the reference.
void Simulator::setPrecision(int p)
{
    ++inputGen;
    if (p < PREC_SINGLE || p > PREC_FP32C)
        throw InvalidArgument("precision must be 1 (single), 2 (double) or 3 (fp32c)");
    if (p != prec) {
        prec = p;
        nPts = 0;
        Felements = 0;
        nStoreMax = 0;
        telements = 0;
        observerInitialized = false;
    }
}

// CLODE::setOpenCL (CLODE.cpp:88-102)
void Simulator::setDevice(int d)
{
    ++inputGen;
    if (d != device) {
        if (stream) {
            hipSetDevice(device);
            hipStreamSynchronize(stream);
            hipEventDestroy(ev0);
            hipEventDestroy(ev1);
            hipStreamDestroy(stream);
            stream = nullptr;
        }
        device = d;
        nPts = 0;
        Felements = 0;
        nStoreMax = 0;
        telements = 0;
        observerInitialized = false;
    }
}

int Simulator::observerId() const
{
    if (kind != 1)
        return 0;
    return observer_map().at(observer);
}

ProgramConfig Simulator::programConfig() const
{
    ProgramConfig c;
    c.prec = prec;
    c.stepper = stepper_map().at(stepper);
    c.observer = observerId();
    c.eVar = kind == 1 ? (int)op.eVarIx : 0;
    c.fVar = kind == 1 ? (int)op.fVarIx : 0;
    c.nVar = prob.nVar;
    c.nPar = prob.nPar;
    c.nAux = prob.nAux;
    c.nWiener = prob.nWiener;
    c.nStore = kind == 1 ? (int)op.maxEventTimestamps : 0;
    c.rhs = rhsSource;
    c.rhs_name = prob.src_file;
    return c;
}

std::shared_ptr<Program> Simulator::currentProgram()
{
    ProgramConfig c = programConfig();
    std::string key = compose_source(c);
    if (!program || key != programKeyCache) {
        program = get_program(c, root.empty() ? std::string() : root);
        programKeyCache = key;
    }
    return program;
}

std::shared_ptr<Program> Simulator::variantProgram(int waves, bool replay)
{
    ProgramConfig c = programConfig();
    c.wavesPerEU = waves;
    c.replay = replay;
    std::string key = compose_source(c);
    auto &slot = programsW[2 * waves + (replay ? 1 : 0)];
    if (!slot.second || key != slot.first) {
        slot.second = get_program(c, root.empty() ? std::string() : root);
        slot.first = key;
    }
    return slot.second;
}

// Kernel for a grid launch of nThreads members.  A kernel whose register
// footprint allows w waves per SIMD leaves each SIMD only w waves to switch
// between while one waits on a dependency.  When the ensemble has at least
// w + 1 waves of members per SIMD and w < 4, the variant compiled for w + 1
// waves per SIMD (registers capped at 512 / (w + 1), the rest spilled) is
// used instead.  Measured gains: lactotroph dopri5 features 1 -> 2 waves,
// 6 % (FP64) and 11 % (fp32c) at 131,072 members, nothing at 65,536; Lorenz
// localmax features on 2^20 members, 2 -> 3 waves in FP64, 6 % (dopri5) and
// 7 % (bs23); 3 -> 4 waves in fp32c, 5 % and 7 %.  Forcing one more wave
// than that (FP64 Lorenz dopri5 at 4 waves) spills enough to lose 26 %.
// Every variant performs the same arithmetic in the same order.
// CLODE_OCCUPANCY_VARIANT=0 in the environment keeps the default kernel.
hipFunction_t Simulator::gridFunction(const char *kernel, unsigned int nThreads, unsigned int block, bool replay)
{
    hipFunction_t fn = (replay ? variantProgram(0, true) : currentProgram())->function(device, kernel);
    lastVariantWaves = 0;
    static const bool enabled = [] {
        const char *e = std::getenv("CLODE_OCCUPANCY_VARIANT");
        return !(e && std::strcmp(e, "0") == 0);
    }();
    if (!enabled || block % 64 != 0)
        return fn;
    int cus = 0, perCU = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
    const size_t simds = (size_t)std::max(cus, 1) * 4;
    hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, (int)block, 0),
              "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    const int waves = (int)((size_t)perCU * block / 64 / 4);  // resident waves per SIMD
    if (waves < 1 || waves >= 4 || (size_t)nThreads < simds * 64 * (size_t)(waves + 1))
        return fn;
    lastVariantWaves = waves + 1;
    return variantProgram(waves + 1, replay)->function(device, kernel);
}

// CLODE::buildCL (CLODE.cpp:155-171) + the subclasses' kernel creation
void Simulator::build()
{
    auto p = currentProgram();
    p->function(device, "clode_transient");
    if (kind == 1) {
        p->function(device, "clode_initialize_observer");
        p->function(device, "clode_features");
    } else if (kind == 2) {
        p->function(device, "clode_trajectory");
    }
    log_msg(LOG_INFO, "built program for device " + std::to_string(device));
}

void Simulator::compileOnly()
{
    currentProgram();
    // with a replay budget, also the program with the record / replay kernels
    if (kind == 1 && replayBudget > 0 && clode_obs_two_pass(observerId()) && prob.nWiener == 0)
        variantProgram(0, true);
}

std::string Simulator::programString()
{
    ProgramConfig c = programConfig();
    std::string s = "// options:";
    for (const auto &o : compile_options(c))
        s += " " + o;
    return s + "\n" + compose_source(c);
}

void Simulator::ensureStream()
{
    set_device(device);
    if (!stream) {
        hip_check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
        hip_check(hipEventCreate(&ev0), "hipEventCreate");
        hip_check(hipEventCreate(&ev1), "hipEventCreate");
    }
}

void Simulator::upload(DeviceBuffer &b, const std::vector<double> &v, size_t elem)
{
    ensureStream();
    if (v.empty())
        return;
    if (elem == 8) {
        hip_check(hipMemcpyAsync(b.ptr, v.data(), v.size() * 8, hipMemcpyHostToDevice, stream), "hipMemcpy H2D");
    } else {
        std::vector<float> f(v.begin(), v.end());
        hip_check(hipMemcpyAsync(b.ptr, f.data(), f.size() * 4, hipMemcpyHostToDevice, stream), "hipMemcpy H2D");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        return;
    }
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

std::vector<double> Simulator::download(const DeviceBuffer &b, size_t n, size_t elem)
{
    ensureStream();
    std::vector<double> out(n);
    if (n == 0 || !b.ptr)
        return out;
    if (elem == 8) {
        hip_check(hipMemcpyAsync(out.data(), b.ptr, n * 8, hipMemcpyDeviceToHost, stream), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    } else {
        std::vector<float> f(n);
        hip_check(hipMemcpyAsync(f.data(), b.ptr, n * 4, hipMemcpyDeviceToHost, stream), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        std::copy(f.begin(), f.end(), out.begin());
    }
    return out;
}

void Simulator::downloadInto(const DeviceBuffer &b, size_t n, size_t elem, double *out)
{
    ensureStream();
    if (n == 0)
        return;
    if (!b.ptr) {
        std::fill(out, out + n, 0.0);
        return;
    }
    if (elem == 8) {
        hip_check(hipMemcpyAsync(out, b.ptr, n * 8, hipMemcpyDeviceToHost, stream), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    } else {
        std::vector<float> f(n);
        hip_check(hipMemcpyAsync(f.data(), b.ptr, n * 4, hipMemcpyDeviceToHost, stream), "hipMemcpy D2H");
        hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
        std::copy(f.begin(), f.end(), out);
    }
}

// CLODE::setNpts (CLODE.cpp:174-242)
void Simulator::setNpts(int newNpts)
{
    ++inputGen;
    if (newNpts <= 0)
        throw InvalidArgument("the ensemble must have at least one member");
    if (newNpts == nPts)
        return;
    set_device(device);
    size_t total = 0, freeb = 0;
    hipMemGetInfo(&freeb, &total);
    size_t largest = (size_t)std::max(prob.nVar, std::max(prob.nPar, prob.nAux)) * (size_t)newNpts * wideSize();
    if (total && largest > total)
        throw InvalidArgument("nPts*nVar, nPts*nPar, or nPts*nAux is too large");
    nPts = newNpts;
    x0.assign((size_t)prob.nVar * nPts, 0.0);
    pars.assign((size_t)prob.nPar * nPts, 0.0);
    rngHost.assign(2 * (size_t)nPts, 0);
    dtHost.assign(nPts, sp.dt);
    d_x0.allocate(device, x0.size() * wideSize());
    d_pars.allocate(device, pars.size() * realSize());
    d_rng.allocate(device, rngHost.size() * 8);
    d_dt.allocate(device, (size_t)nPts * wideSize());
    d_tf.allocate(device, (size_t)nPts * wideSize());
    d_xf.allocate(device, x0.size() * wideSize());
    upload(d_dt, dtHost, wideSize());
    memberTspan = false;
    d_tspan_m.release();
    d_qorder.release();
    d_steps.release();
    seedRNG();
    log_msg(LOG_DEBUG, "set nPts=" + std::to_string(nPts));
}

// CLODE::setProblemData (CLODE.cpp:245-280)
void Simulator::setProblemData(const std::vector<double> &nx0, const std::vector<double> &npars)
{
    ++inputGen;
    if (nx0.size() % prob.nVar != 0) {
        log_msg(LOG_INFO, "Invalid initial condition vector: not a multiple of nVar=" + std::to_string(prob.nVar));
        log_msg(LOG_INFO, "...Initial conditions were not updated!");
        return;
    }
    if (prob.nPar > 0 && npars.size() % prob.nPar != 0) {
        log_msg(LOG_INFO, "Invalid parameter vector: not a multiple of nPar=" + std::to_string(prob.nPar));
        log_msg(LOG_INFO, "...Parameters were not updated!");
        return;
    }
    int nx = (int)(nx0.size() / prob.nVar);
    int np = prob.nPar > 0 ? (int)(npars.size() / prob.nPar) : nx;
    if (nx != np) {
        log_msg(LOG_INFO, "Initial contition and parameter vector dimensions don't match");
        log_msg(LOG_INFO, "...Problem data was not updated!");
        return;
    }
    setNpts(nx);
    setX0(nx0);
    setPars(npars);
}

// CLODE::setX0 (CLODE.cpp:283-315)
void Simulator::setX0(const std::vector<double> &nx0)
{
    ++inputGen;
    if (nPts > 0 && nx0.size() == (size_t)nPts * prob.nVar) {
        x0 = nx0;
        upload(d_x0, x0, wideSize());
    } else {
        log_msg(LOG_INFO, "...Initial conditions were not updated!");
    }
}

// CLODE::setPars (CLODE.cpp:318-350)
void Simulator::setPars(const std::vector<double> &np)
{
    ++inputGen;
    if (nPts > 0 && np.size() == (size_t)nPts * prob.nPar) {
        pars = np;
        upload(d_pars, pars, realSize());
    } else {
        log_msg(LOG_INFO, "Invalid parameter vector: Expected " + std::to_string(nPts) + "*" +
                              std::to_string(prob.nPar) + " elements, received " + std::to_string(np.size()));
        log_msg(LOG_INFO, "...Parameters were not updated!");
    }
}

// CLODE::setTspan (CLODE.cpp:352-375)
void Simulator::setTspan(const std::vector<double> &t)
{
    ++inputGen;
    if (t.size() != 2)
        throw InvalidArgument("tspan must have two elements");
    tspan = t;
}

// CLODE::setSolverParams (CLODE.cpp:377-403): the host dt vector is refilled
// but the device buffer is not (unless dt refresh is enabled)
void Simulator::setSolverParams(const SolverParams &nsp)
{
    ++inputGen;
    if (nsp.nout == 0)
        throw InvalidArgument("nout must be positive");
    sp = nsp;
    std::fill(dtHost.begin(), dtHost.end(), sp.dt);
    if (dtRefresh && nPts > 0)
        upload(d_dt, dtHost, wideSize());
}

// CLODE::seedRNG() (CLODE.cpp:420-444)
void Simulator::seedRNG()
{
    ++inputGen;
    if (nPts <= 0)
        return;
    std::random_device rd;
    std::mt19937_64 gen(rd());
    std::uniform_int_distribution<unsigned long long> dis;
    for (auto &w : rngHost)
        w = dis(gen);
    ensureStream();
    hip_check(hipMemcpyAsync(d_rng.ptr, rngHost.data(), rngHost.size() * 8, hipMemcpyHostToDevice, stream),
              "hipMemcpy H2D");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// CLODE::seedRNG(int) (CLODE.cpp:447-464): word k = seed + k in int arithmetic
void Simulator::seedRNG(int seed)
{
    ++inputGen;
    if (nPts <= 0)
        return;
    for (size_t k = 0; k < rngHost.size(); ++k) {
        int w = (int)((unsigned int)seed + (unsigned int)k);
        rngHost[k] = (unsigned long long)(long long)w;
    }
    ensureStream();
    hip_check(hipMemcpyAsync(d_rng.ptr, rngHost.data(), rngHost.size() * 8, hipMemcpyHostToDevice, stream),
              "hipMemcpy H2D");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

void Simulator::setMemberTspan(const double *t, size_t n)
{
    ++inputGen;
    if (!t) {
        memberTspan = false;
        d_tspan_m.release();
        d_qorder.release();
        return;
    }
    requireData("set_member_tspan");
    if (n != 2 * (size_t)nPts)
        throw InvalidArgument("member tspan must have 2*nPts elements");
    for (size_t i = 0; i < (size_t)nPts; ++i)
        if (!(t[nPts + i] >= t[i]))
            throw InvalidArgument("member tspan: every end must be >= its start");
    // work-queue order: longest span first, so that the longest members start
    // first and the short ones fill the tail
    std::vector<unsigned int> order((size_t)nPts);
    for (size_t i = 0; i < order.size(); ++i)
        order[i] = (unsigned int)i;
    std::stable_sort(order.begin(), order.end(), [&](unsigned int a, unsigned int b) {
        return t[nPts + a] - t[a] > t[nPts + b] - t[b];
    });
    d_tspan_m.allocate(device, n * 8);
    d_qorder.allocate(device, order.size() * 4);
    ensureStream();
    hip_check(hipMemcpyAsync(d_tspan_m.ptr, t, n * 8, hipMemcpyHostToDevice, stream), "hipMemcpy H2D");
    hip_check(hipMemcpyAsync(d_qorder.ptr, order.data(), order.size() * 4, hipMemcpyHostToDevice, stream),
              "hipMemcpy H2D");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    memberTspan = true;
}

void Simulator::requireData(const char *what) const
{
    if (nPts <= 0)
        throw InvalidArgument(std::string(what) + ": problem data (x0, pars) must be set first");
}

ClodeArgs Simulator::makeArgs()
{
    ClodeArgs a;
    std::memset(&a, 0, sizeof(a));
    a.t0 = tspan[0];
    a.t1 = tspan[1];
    a.sp_dt = sp.dt;
    a.sp_dtmax = sp.dtmax;
    a.sp_abstol = sp.abstol;
    a.sp_reltol = sp.reltol;
    a.sp_max_steps = sp.max_steps;
    a.sp_max_store = sp.max_store;
    a.sp_nout = sp.nout;
    a.nPts = (unsigned int)nPts;
    a.op_eVarIx = op.eVarIx;
    a.op_fVarIx = op.fVarIx;
    a.op_maxEventCount = op.maxEventCount;
    a.op_maxEventTimestamps = op.maxEventTimestamps;
    a.op_minXamp = op.minXamp;
    a.op_minIMI = op.minIMI;
    a.op_nHoodRadius = op.nHoodRadius;
    a.op_xUpThresh = op.xUpThresh;
    a.op_xDownThresh = op.xDownThresh;
    a.op_dxUpThresh = op.dxUpThresh;
    a.op_dxDownThresh = op.dxDownThresh;
    a.op_eps_dx = op.eps_dx;
    a.x0 = d_x0.ptr;
    a.pars = d_pars.ptr;
    a.xf = d_xf.ptr;
    a.rng = d_rng.as<unsigned long long>();
    a.dt = d_dt.ptr;
    a.tf = d_tf.ptr;
    a.odr = d_odr.ptr;
    a.odu = d_odu.as<unsigned int>();
    a.F = d_F.ptr;
    a.tr_t = d_t.ptr;
    a.tr_x = d_x.ptr;
    a.tr_dx = d_dx.ptr;
    a.tr_aux = d_aux.ptr;
    a.nStored = d_nStored.as<int>();
    a.tspan_m = memberTspan ? d_tspan_m.ptr : nullptr;
    if (stepCounter) {
        if (!d_steps.ptr || d_steps.bytes != (size_t)nPts * 16) {  // [2][nPts]: steps, trials
            d_steps.allocate(device, (size_t)nPts * 16);
            resetStepCounts();
        }
        a.steps = d_steps.as<unsigned long long>();
    }
    a.q_order = memberTspan ? d_qorder.as<unsigned int>() : nullptr;
    a.q_count = (unsigned int)nPts;
    a.q_stride = 1;
    a.q_slice = queueSlice;
    return a;
}

// Persistent work-queue grid: as many workgroups as can be resident at once
// (occupancy of this kernel at this block size x compute units), but no more
// than the members need.
unsigned int queue_grid(hipFunction_t fn, int device, unsigned int block, size_t members)
{
    int perCU = 0, cus = 0;
    hip_check(hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(&perCU, fn, (int)block, 0),
              "hipModuleOccupancyMaxActiveBlocksPerMultiprocessor");
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "hipDeviceGetAttribute");
    size_t resident = (size_t)std::max(perCU, 1) * (size_t)std::max(cus, 1);
    size_t needed = (members + block - 1) / block;
    return (unsigned int)std::max<size_t>(1, std::min(resident, needed));
}

void Simulator::launch(const char *kernel, const ClodeArgs &a, unsigned int nThreads, bool replay)
{
    auto p = currentProgram();
    ensureStream();
    unsigned int block = blockSize > 0 ? (unsigned int)blockSize : 64u;
    ClodeArgs args = a;
    void *params[] = {&args};
    if (launchMode == 1 && !replay) {
        // work queue: "clode_X" -> "clode_queue_X", counter cleared by a kernel
        std::string qname = std::string("clode_queue_") + (kernel + 6);
        hipFunction_t fn = p->function(device, qname.c_str());
        if (!d_work.ptr)
            d_work.allocate(device, 256);
        args.work_counter = d_work.as<unsigned int>();
        hipFunction_t reset = p->function(device, "clode_queue_reset");
        lastVariantWaves = 0;
        unsigned int one = 1;
        void *rparams[] = {&args.work_counter, &one};
        hip_check(hipModuleLaunchKernel(reset, 1, 1, 1, 64, 1, 1, 0, stream, rparams, nullptr),
                  "launch clode_queue_reset");
        unsigned int grid = queue_grid(fn, device, block, nThreads);
        hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
        hip_check(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, stream, params, nullptr),
                  ("launch " + qname).c_str());
        hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
        hip_check(hipStreamSynchronize(stream), ("kernel " + qname).c_str());
        float ms = 0.f;
        hipEventElapsedTime(&ms, ev0, ev1);
        lastKernelMs = ms;
        return;
    }
    hipFunction_t fn = gridFunction(kernel, nThreads, block, replay);
    unsigned int grid = (nThreads + block - 1) / block;
    hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
    hip_check(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, stream, params, nullptr),
              (std::string("launch ") + kernel).c_str());
    hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
    hip_check(hipStreamSynchronize(stream), (std::string("kernel ") + kernel).c_str());
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev0, ev1);
    lastKernelMs = ms;
}

void Simulator::setStepCounter(bool on)
{
    stepCounter = on;
    if (!on)
        d_steps.release();
}

void Simulator::resetStepCounts()
{
    if (!d_steps.ptr)
        return;
    ensureStream();
    hip_check(hipMemsetAsync(d_steps.ptr, 0, d_steps.bytes, stream), "hipMemset");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

std::vector<unsigned long long> Simulator::getStepCounts(int row)
{
    std::vector<unsigned long long> out(nPts > 0 ? nPts : 0, 0ull);
    if (!d_steps.ptr || out.empty() || d_steps.bytes < (size_t)nPts * 16)
        return out;
    ensureStream();
    hip_check(hipMemcpyAsync(out.data(), d_steps.as<unsigned long long>() + (size_t)row * nPts, out.size() * 8,
                             hipMemcpyDeviceToHost, stream),
              "hipMemcpy D2H");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return out;
}

std::vector<unsigned long long> Simulator::getRngState()
{
    std::vector<unsigned long long> out(rngHost.size(), 0ull);
    if (out.empty() || !d_rng.ptr)
        return out;
    ensureStream();
    hip_check(hipMemcpyAsync(out.data(), d_rng.ptr, out.size() * 8, hipMemcpyDeviceToHost, stream), "hipMemcpy D2H");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return out;
}

// CLODE::transient (CLODE.cpp:468-493)
void Simulator::transient()
{
    ++inputGen;
    requireData("transient");
    launch("clode_transient", makeArgs(), (unsigned int)nPts);
    log_msg(LOG_INFO, "run transient");
}

// CLODE::shiftTspan (CLODE.cpp:495-500)
void Simulator::shiftTspan()
{
    std::vector<double> t = {tspan[1], tspan[1] + (tspan[1] - tspan[0])};
    setTspan(t);
}

// CLODE::shiftX0 (CLODE.cpp:502-515): device-to-device copy xf -> x0
void Simulator::shiftX0()
{
    ++inputGen;
    requireData("shift_x0");
    ensureStream();
    hip_check(hipMemcpyAsync(d_x0.ptr, d_xf.ptr, (size_t)prob.nVar * nPts * wideSize(), hipMemcpyDeviceToDevice,
                             stream),
              "hipMemcpy D2D");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
}

// CLODE::getX0 / getXf / getDt / getTf (CLODE.cpp:517-583)
std::vector<double> Simulator::getX0()
{
    if (nPts <= 0)
        return {};
    x0 = download(d_x0, (size_t)prob.nVar * nPts, wideSize());
    return x0;
}
std::vector<double> Simulator::getXf()
{
    if (nPts <= 0)
        return {};
    return download(d_xf, (size_t)prob.nVar * nPts, wideSize());
}
std::vector<double> Simulator::getDt()
{
    if (nPts <= 0)
        return {};
    return download(d_dt, (size_t)nPts, wideSize());
}
std::vector<double> Simulator::getTf()
{
    if (nPts <= 0)
        return {};
    return download(d_tf, (size_t)nPts, wideSize());
}

// ------------------------------------------------------------------ features

// CLODEfeatures::updateObserverDefineMap (CLODEfeatures.cpp:108-122)
void Simulator::updateObserverInfo()
{
    int obs = observerId();
    int S = (int)op.maxEventTimestamps;
    nFeatures = clode_obs_nfeatures(obs, prob.nVar, prob.nAux, S);
    featureNames = observer_feature_names(obs, prob, S, op.fVarIx);
}

// CLODEfeatures::setObserverParams (CLODEfeatures.cpp:78-106)
void Simulator::setObserverParams(const ObserverParams &nop)
{
    ++inputGen;
    if (nop.eVarIx >= (unsigned int)prob.nVar || nop.fVarIx >= (unsigned int)prob.nVar)
        throw InvalidArgument("observer eVarIx/fVarIx must index a state variable");
    op = nop;
    updateObserverInfo();
}

// CLODEfeatures::setObserver (CLODEfeatures.cpp:63-76)
void Simulator::setObserver(const std::string &name)
{
    ++inputGen;
    if (observer_map().find(name) != observer_map().end()) {
        observer = name;
        updateObserverInfo();
    } else {
        log_msg(LOG_WARN, "unknown observer: " + name + ". Observer method unchanged");
    }
}

// CLODEfeatures::resizeFeaturesVariables (CLODEfeatures.cpp:143-177)
void Simulator::resizeFeaturesVariables()
{
    int obs = observerId();
    int S = (int)op.maxEventTimestamps;
    size_t nr = (size_t)clode_obs_nreal(obs, prob.nVar, prob.nAux, S);
    size_t nu = (size_t)clode_obs_nuint(obs);
    size_t fe = (size_t)nFeatures * nPts;
    if (fe != Felements || nr != odrFields || nu != oduFields || !d_F.ptr) {
        Felements = fe;
        odrFields = nr;
        oduFields = nu;
        d_odr.allocate(device, nr * nPts * wideSize());
        d_odu.allocate(device, nu * nPts * 4);
        d_F.allocate(device, fe * wideSize());
        observerInitialized = false;
    }
}

void Simulator::setReplayBudget(size_t bytes)
{
    replayBudget = bytes;
    replayGen = ~0ull;
    replayCap = 0;
    if (bytes == 0) {
        d_replay.release();
        d_replay_n.release();
    }
}

bool Simulator::replayUsable() const
{
    return kind == 1 && replayBudget > 0 && nPts > 0 && clode_obs_two_pass(observerId()) && prob.nWiener == 0 &&
           launchMode == 0;
}

// Records per member: what the budget holds, at most max_steps (rounded up
// to whole chunks), a multiple of the chunk; the buffer grows when needed and
// is kept.  false (no recording) below one chunk or when the allocation fails.
bool Simulator::prepareReplay()
{
    const size_t rec = (size_t)clode_replay_record(prob.nVar, prob.nAux, prec) * 8;
    const size_t chunk = CLODE_REPLAY_CHUNK;
    size_t cap = replayBudget / ((size_t)nPts * rec);
    const size_t steps = ((size_t)sp.max_steps + chunk - 1) / chunk * chunk;
    cap = std::min(std::min(cap, steps), (size_t)0xfffffff0u) / chunk * chunk;
    replayCap = 0;
    if (cap < chunk)
        return false;
    const size_t bytes = cap * (size_t)nPts * rec;
    try {
        if (d_replay.bytes < bytes || d_replay.device != device)
            d_replay.allocate(device, bytes);
        if (d_replay_n.bytes < (size_t)nPts * 4 || d_replay_n.device != device)
            d_replay_n.allocate(device, (size_t)nPts * 4);
    } catch (const std::exception &e) {
        log_msg(LOG_WARN, std::string("features replay off: ") + e.what());
        d_replay.release();
        d_replay_n.release();
        (void)hipGetLastError();
        return false;
    }
    replayCap = (unsigned int)cap;
    return true;
}

ClodeArgs Simulator::replayArgs()
{
    ClodeArgs a = makeArgs();
    a.replay = d_replay.as<double>();
    a.replay_n = d_replay_n.as<unsigned int>();
    a.replay_cap = replayCap;
    return a;
}

// CLODEfeatures::initializeObserver (CLODEfeatures.cpp:180-212)
void Simulator::initializeObserver()
{
    requireData("initialize_observer");
    resizeFeaturesVariables();
    replayGen = ~0ull;
    if (replayUsable() && prepareReplay()) {
        launch("clode_initialize_observer_record", replayArgs(), (unsigned int)nPts, true);
        replayGen = inputGen;
    } else {
        launch("clode_initialize_observer", makeArgs(), (unsigned int)nPts);
    }
    observerInitialized = true;
}

// CLODEfeatures::features(bool) (CLODEfeatures.cpp:215-220)
void Simulator::features(bool reinit)
{
    observerInitialized = !reinit;
    features();
}

// CLODEfeatures::features() (CLODEfeatures.cpp:222-259)
void Simulator::features()
{
    requireData("features");
    resizeFeaturesVariables();
    if (!observerInitialized)
        initializeObserver();
    lastReplay = replayGen == inputGen && replayUsable() && d_replay.ptr;
    if (lastReplay)
        launch("clode_features_replay", replayArgs(), (unsigned int)nPts, true);
    else
        launch("clode_features", makeArgs(), (unsigned int)nPts);
    ++inputGen;  // xf, dt and tf are stored: the record describes no later run
}

// transient(), shift_x0(), features(true) in one launch: each lane runs its
// member through the three passes (clode_sequence).  Grid launch only; the
// global spans only (per-member spans would give both passes the same span).
void Simulator::transientFeatures(double tr0, double tr1, double fe0, double fe1)
{
    ++inputGen;
    requireData("transient_features");
    if (kind != 1)
        throw InvalidArgument("transient_features needs a feature simulator");
    if (memberTspan)
        throw InvalidArgument("transient_features runs on the global spans; clear the per-member spans first");
    resizeFeaturesVariables();
    ClodeSeqArgs s;
    setTspan({tr0, tr1});
    s.tr = makeArgs();
    setTspan({fe0, fe1});
    s.fe = makeArgs();
    ensureStream();
    const unsigned int n = (unsigned int)nPts;
    const unsigned int block = blockSize > 0 ? (unsigned int)blockSize : 64u;
    hipFunction_t fn = gridFunction("clode_sequence", n, block);
    void *params[] = {&s};
    hip_check(hipEventRecord(ev0, stream), "hipEventRecord");
    hip_check(hipModuleLaunchKernel(fn, (n + block - 1) / block, 1, 1, block, 1, 1, 0, stream, params, nullptr),
              "launch clode_sequence");
    hip_check(hipEventRecord(ev1, stream), "hipEventRecord");
    hip_check(hipStreamSynchronize(stream), "kernel clode_sequence");
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev0, ev1);
    lastKernelMs = ms;
    observerInitialized = true;
}

// CLODEfeatures::getF (CLODEfeatures.cpp:261-275)
std::vector<double> Simulator::getF()
{
    if (nPts <= 0 || !d_F.ptr)
        return std::vector<double>(Felements, 0.0);
    return download(d_F, Felements, wideSize());
}

void Simulator::getFInto(double *out, size_t cap)
{
    if (cap < Felements)
        throw InvalidArgument("output buffer too small: need " + std::to_string(Felements));
    if (nPts <= 0 || !d_F.ptr) {
        std::fill(out, out + Felements, 0.0);
        return;
    }
    downloadInto(d_F, Felements, wideSize(), out);
}

// The Python API's layout, F[member][feature] (features.py:467-507 reshapes
// the reference's feature-major vector to it on the host).
void Simulator::getFMembers(double *out, size_t cap)
{
    if (cap < Felements)
        throw InvalidArgument("output buffer too small: need " + std::to_string(Felements));
    if (nPts <= 0 || !d_F.ptr || Felements == 0) {
        std::fill(out, out + Felements, 0.0);
        return;
    }
    ensureStream();
    if (!d_Fm.ptr || d_Fm.bytes != Felements * 8)
        d_Fm.allocate(device, Felements * 8);
    hipFunction_t fn = currentProgram()->function(device, "clode_f_members");
    const void *src = d_F.ptr;
    void *dst = d_Fm.ptr;
    unsigned int n = (unsigned int)nPts, nF = (unsigned int)(Felements / (size_t)nPts);
    void *params[] = {&src, &dst, &n, &nF};
    hip_check(hipModuleLaunchKernel(fn, (n + 63u) / 64u, (nF + 31u) / 32u, 1, 64, 1, 1, 0, stream, params, nullptr),
              "launch clode_f_members");
    downloadInto(d_Fm, Felements, 8, out);
}

// ---------------------------------------------------------------- trajectory

// CLODEtrajectory::resizeTrajectoryVariables (CLODEtrajectory.cpp:45-94)
void Simulator::resizeTrajectoryVariables()
{
    unsigned int store = sp.max_store;
    size_t largest = (size_t)std::max(1, std::max(prob.nVar, prob.nAux)) * nPts * store * realSize();
    size_t freeb = 0, total = 0;
    set_device(device);
    hipMemGetInfo(&freeb, &total);
    if (total && largest > total)
        throw InvalidArgument("nPts*nStoreMax*nVar*realSize or nPts*nStoreMax*nAux*realSize is too big");
    size_t te = (size_t)store * nPts;
    if (nStoreMax != store || telements != te || !d_t.ptr) {
        nStoreMax = store;
        telements = te;
        xelements = (size_t)prob.nVar * te;
        auxelements = prob.nAux > 0 ? (size_t)prob.nAux * te : 1;
        d_t.allocate(device, te * wideSize());
        d_x.allocate(device, xelements * realSize());
        d_dx.allocate(device, xelements * realSize());
        d_aux.allocate(device, auxelements * realSize());
        d_nStored.allocate(device, (size_t)nPts * 4);
    }
}

// CLODEtrajectory::trajectory (CLODEtrajectory.cpp:97-130)
void Simulator::trajectory()
{
    ++inputGen;
    requireData("trajectory");
    resizeTrajectoryVariables();
    launch("clode_trajectory", makeArgs(), (unsigned int)nPts);
}

// CLODEtrajectory::getT / getX / getDx / getAux / getNstored (CLODEtrajectory.cpp:132-205)
std::vector<double> Simulator::getT() { return download(d_t, telements, wideSize()); }
std::vector<double> Simulator::getX() { return download(d_x, xelements, realSize()); }
std::vector<double> Simulator::getDx() { return download(d_dx, xelements, realSize()); }
std::vector<double> Simulator::getAux() { return download(d_aux, auxelements, realSize()); }
std::vector<int> Simulator::getNstored()
{
    std::vector<int> out(nPts > 0 ? nPts : 0);
    if (!d_nStored.ptr || out.empty())
        return out;
    ensureStream();
    hip_check(hipMemcpyAsync(out.data(), d_nStored.ptr, out.size() * 4, hipMemcpyDeviceToHost, stream),
              "hipMemcpy D2H");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    return out;
}

}  // namespace clode
