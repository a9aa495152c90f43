// extern "C" entry points declared in include/clode_hip.h.  Each wraps one
// Simulator member; exceptions become status codes plus a thread-local
// message (InvalidArgument -> CLODE_ERR_INVALID, everything else ->
// CLODE_ERR_RUNTIME).
#include "../../../include/clode_hip.h"

#include <cstring>

#include "batch.h"
#include "simulator.h"

using namespace clode;

struct clode_sim {
    Simulator *s;
};

struct clode_batch {
    Batch *b;
};

static thread_local std::string g_last_error;

#define CLODE_TRY try {
#define CLODE_CATCH                                  \
    }                                                \
    catch (const InvalidArgument &e)                 \
    {                                                \
        g_last_error = e.what();                     \
        return CLODE_ERR_INVALID;                    \
    }                                                \
    catch (const std::out_of_range &e)               \
    {                                                \
        g_last_error = e.what();                     \
        return CLODE_ERR_INVALID;                    \
    }                                                \
    catch (const std::exception &e)                  \
    {                                                \
        g_last_error = e.what();                     \
        return CLODE_ERR_RUNTIME;                    \
    }                                                \
    catch (...)                                      \
    {                                                \
        g_last_error = "unknown error";              \
        return CLODE_ERR_RUNTIME;                    \
    }                                                \
    return CLODE_OK;

#define CHECK_SIM(sim)                                        \
    if (!(sim) || !(sim)->s) {                                \
        g_last_error = "null simulator handle";               \
        return CLODE_ERR_INVALID;                             \
    }

#define CHECK_BATCH(batch)                                    \
    if (!(batch) || !(batch)->b) {                            \
        g_last_error = "null batch handle";                   \
        return CLODE_ERR_INVALID;                             \
    }

// a caller's pointer that must not be null
template <class T>
static T *need(T *p, const char *what)
{
    if (!p)
        throw InvalidArgument(std::string("null ") + what);
    return p;
}

// a caller's input array of n elements (null only when n == 0)
static std::vector<double> in_vec(const double *p, size_t n, const char *what)
{
    if (!p && n > 0)
        throw InvalidArgument(std::string("null ") + what);
    return p ? std::vector<double>(p, p + n) : std::vector<double>();
}

static std::vector<std::string> split_names(const char *s)
{
    std::vector<std::string> out;
    if (!s || !*s)
        return out;
    std::string cur;
    for (const char *p = s; *p; ++p) {
        if (*p == '\n') {
            out.push_back(cur);
            cur.clear();
        } else {
            cur += *p;
        }
    }
    out.push_back(cur);
    return out;
}

static ProblemInfo to_pi(const clode_problem_info *pi)
{
    if (!pi || !pi->src_file)
        throw InvalidArgument("problem info with a source file is required");
    ProblemInfo p;
    p.src_file = pi->src_file;
    p.nVar = pi->num_var;
    p.nPar = pi->num_par;
    p.nAux = pi->num_aux;
    p.nWiener = pi->num_noise;
    p.varNames = split_names(pi->var_names);
    p.parNames = split_names(pi->par_names);
    p.auxNames = split_names(pi->aux_names);
    return p;
}

static ObserverParams to_op(const clode_observer_params *o)
{
    ObserverParams p;
    if (!o)
        return p;
    p.eVarIx = o->eVarIx;
    p.fVarIx = o->fVarIx;
    p.maxEventCount = o->maxEventCount;
    p.maxEventTimestamps = o->maxEventTimestamps;
    p.minXamp = o->minXamp;
    p.minIMI = o->minIMI;
    p.nHoodRadius = o->nHoodRadius;
    p.xUpThresh = o->xUpThresh;
    p.xDownThresh = o->xDownThresh;
    p.dxUpThresh = o->dxUpThresh;
    p.dxDownThresh = o->dxDownThresh;
    p.eps_dx = o->eps_dx;
    return p;
}

static int copy_string(const std::string &s, char *buf, size_t cap, size_t *needed)
{
    if (needed)
        *needed = s.size() + 1;
    if (buf && cap > 0) {
        size_t n = std::min(cap - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    }
    return CLODE_OK;
}

static std::string join(const std::vector<std::string> &v)
{
    std::string s;
    for (size_t i = 0; i < v.size(); ++i) {
        if (i)
            s += '\n';
        s += v[i];
    }
    return s;
}

static int copy_vec(const std::vector<double> &v, double *out, size_t n)
{
    if (!out && n > 0) {
        g_last_error = "null output buffer";
        return CLODE_ERR_INVALID;
    }
    if (n < v.size()) {
        g_last_error = "output buffer too small: need " + std::to_string(v.size());
        return CLODE_ERR_INVALID;
    }
    std::copy(v.begin(), v.end(), out);
    return CLODE_OK;
}

extern "C" {

const char *clode_last_error(void) { return g_last_error.c_str(); }
const char *clode_version(void) { return "clode_amd 0.1.0 (gfx950, hipRTC)"; }

int clode_device_count(int *count)
{
    CLODE_TRY
    *need(count, "count") = device_count();
    CLODE_CATCH
}

int clode_device_query(int device, clode_device_info *info)
{
    CLODE_TRY
    if (!info)
        throw InvalidArgument("null info");
    int n = device_count();
    if (device < 0 || device >= n)
        throw InvalidArgument("device ordinal out of range");
    hipDeviceProp_t p;
    hip_check(hipGetDeviceProperties(&p, device), "hipGetDeviceProperties");
    std::memset(info, 0, sizeof(*info));
    std::strncpy(info->name, p.name, sizeof(info->name) - 1);
    std::strncpy(info->arch, p.gcnArchName, sizeof(info->arch) - 1);
    info->compute_units = p.multiProcessorCount;
    info->max_clock_khz = p.clockRate;
    info->max_work_group_size = p.maxThreadsPerBlock;
    info->device_memory_size = p.totalGlobalMem;
    info->max_memory_alloc_size = p.totalGlobalMem;
    info->double_support = 1;
    CLODE_CATCH
}

int clode_create(clode_sim **out, int kind, const clode_problem_info *pi, const char *stepper, const char *observer,
                 const clode_observer_params *op, int precision, int device, const char *clode_root)
{
    CLODE_TRY
    if (!out)
        throw InvalidArgument("null output handle");
    *out = nullptr;
    if (kind < CLODE_KIND_BASE || kind > CLODE_KIND_TRAJECTORY)
        throw InvalidArgument("unknown simulator kind");
    Simulator *s = new Simulator(kind, to_pi(pi), stepper ? stepper : "rk4", observer ? observer : "",
                                 to_op(op), precision, device, clode_root ? clode_root : "");
    *out = new clode_sim{s};
    CLODE_CATCH
}

int clode_destroy(clode_sim *sim)
{
    CLODE_TRY
    if (sim) {
        delete sim->s;
        delete sim;
    }
    CLODE_CATCH
}

int clode_set_problem_info(clode_sim *sim, const clode_problem_info *pi)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setProblemInfo(to_pi(pi));
    CLODE_CATCH
}
int clode_set_stepper(clode_sim *sim, const char *stepper)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setStepper(stepper ? stepper : "");
    CLODE_CATCH
}
int clode_set_precision(clode_sim *sim, int precision)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setPrecision(precision);
    CLODE_CATCH
}
int clode_set_device(clode_sim *sim, int device)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setDevice(device);
    CLODE_CATCH
}
int clode_build(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->build();
    CLODE_CATCH
}
int clode_compile(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->compileOnly();
    CLODE_CATCH
}

int clode_set_problem_data(clode_sim *sim, const double *x0, size_t nx0, const double *pars, size_t npars)
{
    CHECK_SIM(sim) CLODE_TRY
    sim->s->setProblemData(in_vec(x0, nx0, "x0"), in_vec(pars, npars, "pars"));
    CLODE_CATCH
}
int clode_set_tspan(clode_sim *sim, const double *tspan, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setTspan(in_vec(tspan, n, "tspan"));
    CLODE_CATCH
}
int clode_set_x0(clode_sim *sim, const double *x0, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setX0(in_vec(x0, n, "x0"));
    CLODE_CATCH
}
int clode_set_pars(clode_sim *sim, const double *pars, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setPars(in_vec(pars, n, "pars"));
    CLODE_CATCH
}
int clode_set_solver_params(clode_sim *sim, const clode_solver_params *sp)
{
    CHECK_SIM(sim) CLODE_TRY
    if (!sp)
        throw InvalidArgument("null solver params");
    SolverParams p;
    p.dt = sp->dt;
    p.dtmax = sp->dtmax;
    p.abstol = sp->abstol;
    p.reltol = sp->reltol;
    p.max_steps = sp->max_steps;
    p.max_store = sp->max_store;
    p.nout = sp->nout;
    sim->s->setSolverParams(p);
    CLODE_CATCH
}
int clode_seed_rng(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->seedRNG();
    CLODE_CATCH
}
int clode_seed_rng_int(clode_sim *sim, int seed)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->seedRNG(seed);
    CLODE_CATCH
}

int clode_transient(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->transient();
    CLODE_CATCH
}
int clode_shift_tspan(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->shiftTspan();
    CLODE_CATCH
}
int clode_shift_x0(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->shiftX0();
    CLODE_CATCH
}

int clode_get_n_pts(clode_sim *sim, int *n)
{
    CHECK_SIM(sim) CLODE_TRY *need(n, "n") = sim->s->nPts;
    CLODE_CATCH
}
int clode_get_tspan(clode_sim *sim, double *t)
{
    CHECK_SIM(sim) CLODE_TRY need(t, "tspan output")[0] = sim->s->tspan[0];
    t[1] = sim->s->tspan[1];
    CLODE_CATCH
}
int clode_get_solver_params(clode_sim *sim, clode_solver_params *sp)
{
    CHECK_SIM(sim) CLODE_TRY
    const SolverParams &p = sim->s->sp;
    need(sp, "solver params output");
    sp->dt = p.dt;
    sp->dtmax = p.dtmax;
    sp->abstol = p.abstol;
    sp->reltol = p.reltol;
    sp->max_steps = p.max_steps;
    sp->max_store = p.max_store;
    sp->nout = p.nout;
    CLODE_CATCH
}
int clode_get_pars(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->pars, out, n);
    CLODE_CATCH
}
int clode_get_x0(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->getX0(), out, n);
    CLODE_CATCH
}
int clode_get_xf(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->getXf(), out, n);
    CLODE_CATCH
}
int clode_get_dt(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->getDt(), out, n);
    CLODE_CATCH
}
int clode_get_tf(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->getTf(), out, n);
    CLODE_CATCH
}
int clode_get_available_steppers(clode_sim *sim, char *buf, size_t cap, size_t *needed)
{
    CLODE_TRY
    std::vector<std::string> v;
    for (const auto &kv : stepper_map())
        v.push_back(kv.first);
    return copy_string(join(v), buf, cap, needed);
    CLODE_CATCH
}
int clode_get_program_string(clode_sim *sim, char *buf, size_t cap, size_t *needed)
{
    CHECK_SIM(sim) CLODE_TRY return copy_string(sim->s->programString(), buf, cap, needed);
    CLODE_CATCH
}
int clode_get_problem_info(clode_sim *sim, int *nvar, int *npar, int *naux, int *nnoise)
{
    CHECK_SIM(sim) CLODE_TRY
    *need(nvar, "nvar") = sim->s->prob.nVar;
    *need(npar, "npar") = sim->s->prob.nPar;
    *need(naux, "naux") = sim->s->prob.nAux;
    *need(nnoise, "nnoise") = sim->s->prob.nWiener;
    CLODE_CATCH
}
int clode_get_precision(clode_sim *sim, int *precision)
{
    CHECK_SIM(sim) CLODE_TRY *need(precision, "precision") = sim->s->prec;
    CLODE_CATCH
}
int clode_get_stepper(clode_sim *sim, char *buf, size_t cap, size_t *needed)
{
    CHECK_SIM(sim) CLODE_TRY return copy_string(sim->s->stepper, buf, cap, needed);
    CLODE_CATCH
}
int clode_print_status(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY
    Simulator &s = *sim->s;
    log_msg(LOG_INFO, "------------------");
    log_msg(LOG_INFO, "   " + s.prob.src_file);
    log_msg(LOG_INFO, "   nVar=" + std::to_string(s.prob.nVar));
    log_msg(LOG_INFO, "   nPar=" + std::to_string(s.prob.nPar));
    log_msg(LOG_INFO, "   nAux=" + std::to_string(s.prob.nAux));
    log_msg(LOG_INFO, "   nWiener=" + std::to_string(s.prob.nWiener));
    log_msg(LOG_INFO, std::string("Using ") +
                          (s.prec == PREC_SINGLE ? "single" : s.prec == PREC_DOUBLE ? "double" : "fp32-compensated") +
                          " precision.");
    log_msg(LOG_INFO, "Using stepper: " + s.stepper);
    CLODE_CATCH
}

int clode_set_observer_params(clode_sim *sim, const clode_observer_params *op)
{
    CHECK_SIM(sim) CLODE_TRY
    if (!op)
        throw InvalidArgument("null observer params");
    sim->s->setObserverParams(to_op(op));
    CLODE_CATCH
}
int clode_set_observer(clode_sim *sim, const char *observer)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setObserver(observer ? observer : "");
    CLODE_CATCH
}
int clode_initialize_observer(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->initializeObserver();
    CLODE_CATCH
}
int clode_is_observer_initialized(clode_sim *sim, int *initialized)
{
    CHECK_SIM(sim) CLODE_TRY *need(initialized, "initialized") = sim->s->observerInitialized ? 1 : 0;
    CLODE_CATCH
}
int clode_features(clode_sim *sim, int reinitialize)
{
    CHECK_SIM(sim) CLODE_TRY
    if (sim->s->kind != CLODE_KIND_FEATURES)
        throw InvalidArgument("features() needs a feature simulator");
    if (reinitialize < 0)
        sim->s->features();
    else
        sim->s->features(reinitialize != 0);
    CLODE_CATCH
}
int clode_get_observer_params(clode_sim *sim, clode_observer_params *o)
{
    CHECK_SIM(sim) CLODE_TRY
    const ObserverParams &p = sim->s->op;
    need(o, "observer params output");
    o->eVarIx = p.eVarIx;
    o->fVarIx = p.fVarIx;
    o->maxEventCount = p.maxEventCount;
    o->maxEventTimestamps = p.maxEventTimestamps;
    o->minXamp = p.minXamp;
    o->minIMI = p.minIMI;
    o->nHoodRadius = p.nHoodRadius;
    o->xUpThresh = p.xUpThresh;
    o->xDownThresh = p.xDownThresh;
    o->dxUpThresh = p.dxUpThresh;
    o->dxDownThresh = p.dxDownThresh;
    o->eps_dx = p.eps_dx;
    CLODE_CATCH
}
int clode_get_observer_name(clode_sim *sim, char *buf, size_t cap, size_t *needed)
{
    CHECK_SIM(sim) CLODE_TRY return copy_string(sim->s->observer, buf, cap, needed);
    CLODE_CATCH
}
int clode_get_n_features(clode_sim *sim, int *n)
{
    CHECK_SIM(sim) CLODE_TRY *need(n, "n") = sim->s->nFeatures;
    CLODE_CATCH
}
int clode_get_f(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY if (!out && n > 0) throw clode::InvalidArgument("null output buffer");
    sim->s->getFInto(out, n);
    CLODE_CATCH
}
int clode_get_f_members(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY if (!out && n > 0) throw clode::InvalidArgument("null output buffer");
    sim->s->getFMembers(out, n);
    CLODE_CATCH
}
int clode_get_feature_names(clode_sim *sim, char *buf, size_t cap, size_t *needed)
{
    CHECK_SIM(sim) CLODE_TRY return copy_string(join(sim->s->featureNames), buf, cap, needed);
    CLODE_CATCH
}
int clode_get_available_observers(clode_sim *sim, char *buf, size_t cap, size_t *needed)
{
    CLODE_TRY
    std::vector<std::string> v;
    for (const auto &kv : observer_map())
        v.push_back(kv.first);
    return copy_string(join(v), buf, cap, needed);
    CLODE_CATCH
}

int clode_trajectory(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY
    if (sim->s->kind != CLODE_KIND_TRAJECTORY)
        throw InvalidArgument("trajectory() needs a trajectory simulator");
    sim->s->trajectory();
    CLODE_CATCH
}
int clode_get_trajectory_sizes(clode_sim *sim, size_t *nt, size_t *nx, size_t *naux)
{
    CHECK_SIM(sim) CLODE_TRY
    *need(nt, "nt") = sim->s->telements;
    *need(nx, "nx") = sim->s->xelements;
    *need(naux, "naux") = sim->s->auxelements;
    CLODE_CATCH
}
int clode_get_t(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->getT(), out, n);
    CLODE_CATCH
}
int clode_get_x(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->getX(), out, n);
    CLODE_CATCH
}
int clode_get_dx(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->getDx(), out, n);
    CLODE_CATCH
}
int clode_get_aux(clode_sim *sim, double *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY return copy_vec(sim->s->getAux(), out, n);
    CLODE_CATCH
}
int clode_get_n_stored(clode_sim *sim, int *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY
    std::vector<int> v = sim->s->getNstored();
    if (n < v.size())
        throw InvalidArgument("output buffer too small");
    if (!v.empty())
        need(out, "output buffer");
    std::copy(v.begin(), v.end(), out);
    CLODE_CATCH
}

int clode_set_log_level(int level)
{
    CLODE_TRY set_log_level(level);
    CLODE_CATCH
}
int clode_get_log_level(int *level)
{
    CLODE_TRY *need(level, "level") = get_log_level();
    CLODE_CATCH
}

int clode_set_queue_slice(clode_sim *sim, int slice)
{
    CHECK_SIM(sim) CLODE_TRY
    if (slice < 1)
        throw InvalidArgument("queue slice must be at least 1");
    sim->s->queueSlice = (unsigned int)slice;
    CLODE_CATCH
}

int clode_batch_create(clode_batch **out, clode_sim *const *sims, size_t n)
{
    if (!out) {
        g_last_error = "null output handle";
        return CLODE_ERR_INVALID;
    }
    *out = nullptr;
    CLODE_TRY
    if (!sims)
        throw InvalidArgument("null simulator array");
    std::vector<Simulator *> v;
    for (size_t k = 0; k < n; ++k) {
        if (!sims[k] || !sims[k]->s)
            throw InvalidArgument("null simulator handle in batch");
        v.push_back(sims[k]->s);
    }
    auto *h = new clode_batch{nullptr};
    try {
        h->b = new Batch(v);
    } catch (...) {
        delete h;
        throw;
    }
    *out = h;
    CLODE_CATCH
}
int clode_batch_compile(clode_sim *const *sims, size_t n)
{
    CLODE_TRY
    if (!sims)
        throw InvalidArgument("null simulator array");
    std::vector<Simulator *> v;
    for (size_t k = 0; k < n; ++k) {
        if (!sims[k] || !sims[k]->s)
            throw InvalidArgument("null simulator handle in batch");
        v.push_back(sims[k]->s);
    }
    Batch::compileOnly(v);
    CLODE_CATCH
}
int clode_batch_destroy(clode_batch *batch)
{
    if (!batch)
        return CLODE_OK;
    CLODE_TRY
    delete batch->b;
    delete batch;
    CLODE_CATCH
}
int clode_batch_set_launch(clode_batch *batch, int block_size, int slice)
{
    CHECK_BATCH(batch) CLODE_TRY batch->b->setLaunch(block_size, slice);
    CLODE_CATCH
}
int clode_batch_get_launch(clode_batch *batch, int *block_size, int *slice)
{
    CHECK_BATCH(batch) CLODE_TRY
    *need(block_size, "block_size") = batch->b->block();
    *need(slice, "slice") = batch->b->slice();
    CLODE_CATCH
}
int clode_batch_transient(clode_batch *batch)
{
    CHECK_BATCH(batch) CLODE_TRY batch->b->transient();
    CLODE_CATCH
}
int clode_batch_initialize_observer(clode_batch *batch)
{
    CHECK_BATCH(batch) CLODE_TRY batch->b->initializeObserver();
    CLODE_CATCH
}
int clode_batch_features(clode_batch *batch, int reinitialize)
{
    CHECK_BATCH(batch) CLODE_TRY
    if (reinitialize < -1 || reinitialize > 1)
        throw InvalidArgument("reinitialize must be -1, 0 or 1");
    batch->b->features(reinitialize);
    CLODE_CATCH
}
int clode_batch_shift_x0(clode_batch *batch)
{
    CHECK_BATCH(batch) CLODE_TRY batch->b->shiftX0();
    CLODE_CATCH
}
int clode_batch_sweep(clode_batch *batch, int rounds, int use_graph)
{
    CHECK_BATCH(batch) CLODE_TRY batch->b->sweep(rounds, use_graph != 0);
    CLODE_CATCH
}
int clode_batch_autotune(clode_batch *batch, int stride, unsigned int max_steps, int *block_size, int *slice)
{
    CHECK_BATCH(batch) CLODE_TRY
    int b = 0, s = 0;
    batch->b->autotune(stride, max_steps, b, s);
    if (block_size)
        *block_size = b;
    if (slice)
        *slice = s;
    CLODE_CATCH
}
int clode_batch_autotune_table(clode_batch *batch, double *rows, int capacity, int *n_rows, int *from_cache)
{
    CHECK_BATCH(batch) CLODE_TRY
    const auto &t = batch->b->tuneTable;
    if (capacity < 0 || (capacity > 0 && !rows))
        throw clode::InvalidArgument("autotune table needs capacity >= 0 and a row buffer");
    const int n = std::min(capacity, (int)t.size());
    for (int r = 0; r < n; ++r) {
        rows[5 * r + 0] = t[r].block;
        rows[5 * r + 1] = t[r].slice;
        rows[5 * r + 2] = t[r].median;
        rows[5 * r + 3] = t[r].min;
        rows[5 * r + 4] = t[r].max;
    }
    *need(n_rows, "n_rows") = (int)t.size();
    if (from_cache)
        *from_cache = batch->b->tuneFromCache ? 1 : 0;
    CLODE_CATCH
}
int clode_batch_last_ms(clode_batch *batch, double *ms)
{
    CHECK_BATCH(batch) CLODE_TRY *need(ms, "ms") = batch->b->lastMs;
    CLODE_CATCH
}

int clode_set_launch(clode_sim *sim, int mode, int block_size)
{
    CHECK_SIM(sim) CLODE_TRY
    if (mode < 0 || mode > 1)
        throw InvalidArgument("launch mode must be 0 (grid) or 1 (persistent)");
    if (block_size < 0 || block_size > 256 || (block_size % 64) != 0)
        throw InvalidArgument("block size must be 0 (auto) or a multiple of 64 up to 256");
    sim->s->launchMode = mode;
    sim->s->blockSize = block_size;
    CLODE_CATCH
}
int clode_set_member_tspan(clode_sim *sim, const double *tspan_m, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setMemberTspan(tspan_m, n);
    CLODE_CATCH
}
int clode_set_dt_refresh(clode_sim *sim, int enable)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->dtRefresh = enable != 0;
    CLODE_CATCH
}
int clode_last_kernel_ms(clode_sim *sim, double *ms)
{
    CHECK_SIM(sim) CLODE_TRY *need(ms, "ms") = sim->s->lastKernelMs;
    CLODE_CATCH
}
int clode_last_launch_variant(clode_sim *sim, int *waves_per_simd)
{
    CHECK_SIM(sim) CLODE_TRY *need(waves_per_simd, "waves_per_simd") = sim->s->lastVariantWaves;
    CLODE_CATCH
}
int clode_set_replay_budget(clode_sim *sim, unsigned long long bytes)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setReplayBudget((size_t)bytes);
    CLODE_CATCH
}
int clode_get_replay_info(clode_sim *sim, unsigned int *records_per_member, int *replayed)
{
    CHECK_SIM(sim) CLODE_TRY
    *need(records_per_member, "records_per_member") = sim->s->replayCap;
    *need(replayed, "replayed") = sim->s->lastReplay ? 1 : 0;
    CLODE_CATCH
}
int clode_set_step_counter(clode_sim *sim, int enable)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->setStepCounter(enable != 0);
    CLODE_CATCH
}
int clode_reset_step_counts(clode_sim *sim)
{
    CHECK_SIM(sim) CLODE_TRY sim->s->resetStepCounts();
    CLODE_CATCH
}
int clode_get_step_counts(clode_sim *sim, unsigned long long *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY
    std::vector<unsigned long long> v = sim->s->getStepCounts();
    if (n < v.size())
        throw InvalidArgument("output buffer too small");
    if (!v.empty())
        need(out, "output buffer");
    std::copy(v.begin(), v.end(), out);
    CLODE_CATCH
}
int clode_transient_features(clode_sim *sim, const double *transient_tspan2, const double *features_tspan2)
{
    CHECK_SIM(sim) CLODE_TRY
    need(transient_tspan2, "transient span");
    need(features_tspan2, "features span");
    sim->s->transientFeatures(transient_tspan2[0], transient_tspan2[1], features_tspan2[0], features_tspan2[1]);
    CLODE_CATCH
}
int clode_get_trial_counts(clode_sim *sim, unsigned long long *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY
    std::vector<unsigned long long> v = sim->s->getStepCounts(1);
    if (n < v.size())
        throw InvalidArgument("output buffer too small");
    if (!v.empty())
        need(out, "output buffer");
    std::copy(v.begin(), v.end(), out);
    CLODE_CATCH
}
int clode_get_rng_state(clode_sim *sim, unsigned long long *out, size_t n)
{
    CHECK_SIM(sim) CLODE_TRY
    std::vector<unsigned long long> v = sim->s->getRngState();
    if (n < v.size())
        throw InvalidArgument("output buffer too small");
    if (!v.empty())
        need(out, "output buffer");
    std::copy(v.begin(), v.end(), out);
    CLODE_CATCH
}

}  // extern "C"
