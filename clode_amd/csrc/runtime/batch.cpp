// Batched work-queue launches over several simulators (see batch.h).
#include "batch.h"

#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>

namespace clode {

Batch::Batch(const std::vector<Simulator *> &sims) : sims_(sims)
{
    if (sims_.empty() || sims_.size() > CLODE_BATCH_MAX)
        throw InvalidArgument("a batch holds 1 to " + std::to_string(CLODE_BATCH_MAX) + " simulators");
    for (Simulator *s : sims_)
        if (!s)
            throw InvalidArgument("null simulator in batch");
    for (size_t i = 0; i < sims_.size(); ++i)
        for (size_t j = i + 1; j < sims_.size(); ++j)
            if (sims_[i] == sims_[j])
                throw InvalidArgument("a simulator appears twice in the batch");
    device_ = sims_[0]->device;
    prec_ = sims_[0]->prec;
    features_ = sims_[0]->kind == 1;
    for (Simulator *s : sims_) {
        if (s->kind == 2)
            throw InvalidArgument("trajectory simulators cannot be batched");
        if ((s->kind == 1) != features_)
            throw InvalidArgument("a batch holds feature simulators only, or base simulators only");
    }
    set_device(device_);
    hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
    hip_check(hipEventCreate(&ev0_), "hipEventCreate");
    hip_check(hipEventCreate(&ev1_), "hipEventCreate");
    counters_.allocate(device_, 256);
}

Batch::~Batch()
{
    if (stream_) {
        hipSetDevice(device_);
        hipStreamSynchronize(stream_);
        hipEventDestroy(ev0_);
        hipEventDestroy(ev1_);
        hipStreamDestroy(stream_);
    }
}

void Batch::setLaunch(int block, int slice)
{
    if (block < 64 || block > 256 || block % 64 != 0)
        throw InvalidArgument("batch block size must be 64, 128, 192 or 256");
    if (slice < 1)
        throw InvalidArgument("batch slice must be at least 1");
    block_ = block;
    slice_ = slice;
}

void Batch::compileOnly(const std::vector<Simulator *> &sims)
{
    std::vector<ProgramConfig> cfgs;
    for (Simulator *s : sims) {
        if (!s)
            throw InvalidArgument("null simulator in batch");
        cfgs.push_back(s->programConfig());
    }
    std::string src = compose_batch_source(cfgs);
    get_program_source(src, compile_options(cfgs[0]), sims[0]->root);
}

std::shared_ptr<Program> Batch::program()
{
    std::vector<ProgramConfig> cfgs;
    for (Simulator *s : sims_)
        cfgs.push_back(s->programConfig());
    std::string src = compose_batch_source(cfgs);
    if (!prog_ || src != progKey_) {
        prog_ = get_program_source(src, compile_options(cfgs[0]), sims_[0]->root);
        progKey_ = src;
    }
    return prog_;
}

// every member simulator has data, lives on the batch's device and has
// finished its own queued work; feature buffers are sized
void Batch::prepare()
{
    for (Simulator *s : sims_) {
        if (s->device != device_)
            throw InvalidArgument("all simulators of a batch must use the same device");
        if (s->prec != prec_)
            throw InvalidArgument("all simulators of a batch must use the same precision");
        s->requireData("batch");
        s->ensureStream();
        ++s->inputGen;  // batch launches write the member buffers: no replay of an earlier record
        if (features_)
            s->resizeFeaturesVariables();
        s->makeArgs();  // allocates the step counters if enabled, outside any capture
        hip_check(hipStreamSynchronize(s->stream), "hipStreamSynchronize");
    }
    set_device(device_);
}

// resident workgroups of fn at the current block size (cached: the query is
// made before any graph capture begins)
unsigned int Batch::gridFor(hipFunction_t fn, size_t members)
{
    auto key = std::make_pair((void *)fn, block_);
    auto it = resident_.find(key);
    if (it == resident_.end())
        it = resident_.emplace(key, queue_grid(fn, device_, (unsigned int)block_, (size_t)1 << 40)).first;
    size_t needed = (members + block_ - 1) / block_;
    return (unsigned int)std::max<size_t>(1, std::min<size_t>(it->second, needed));
}

std::vector<ClodeArgs> Batch::args()
{
    std::vector<ClodeArgs> a;
    for (size_t k = 0; k < sims_.size(); ++k) {
        ClodeArgs x = sims_[k]->makeArgs();
        x.work_counter = counters_.as<unsigned int>() + k;
        x.q_slice = (unsigned int)slice_;
        a.push_back(x);
    }
    return a;
}

// clear the counters, then launch the dispatching kernel on the batch stream
void Batch::enqueue(const char *kernel, std::vector<ClodeArgs> a)
{
    auto p = program();
    hipFunction_t reset = p->function(device_, "clode_queue_reset");
    hipFunction_t fn = p->function(device_, kernel);
    unsigned int *ctr = counters_.as<unsigned int>();
    unsigned int n = (unsigned int)a.size();
    void *rparams[] = {&ctr, &n};
    hip_check(hipModuleLaunchKernel(reset, 1, 1, 1, 64, 1, 1, 0, stream_, rparams, nullptr),
              "launch clode_queue_reset");
    size_t members = 0;
    for (const auto &x : a)
        members += x.q_count;
    if (members == 0)
        return;
    unsigned int grid = gridFor(fn, members);
    void *params[] = {a.data()};
    hip_check(hipModuleLaunchKernel(fn, grid, 1, 1, (unsigned int)block_, 1, 1, 0, stream_, params, nullptr),
              (std::string("launch ") + kernel).c_str());
}

void Batch::enqueueShift(std::vector<ClodeArgs> a)
{
    auto p = program();
    hipFunction_t fn = p->function(device_, "clode_batch_shift");
    size_t elems = 0;
    for (size_t k = 0; k < sims_.size(); ++k)
        elems = std::max(elems, (size_t)sims_[k]->prob.nVar * sims_[k]->nPts);
    unsigned int grid = (unsigned int)std::min<size_t>(2048, (elems + 255) / 256);
    void *params[] = {a.data()};
    hip_check(hipModuleLaunchKernel(fn, std::max(grid, 1u), 1, 1, 256, 1, 1, 0, stream_, params, nullptr),
              "launch clode_batch_shift");
}

// CLODEfeatures::features semantics per simulator (CLODEfeatures.cpp:215-259):
// reinit 1 re-initialises, 0 continues as if initialised, -1 initialises only
// where needed; the observer initialisation runs as one batched launch over
// the simulators that need it
void Batch::enqueueFeatures(int reinit, std::vector<bool> &initialised)
{
    std::vector<ClodeArgs> init = args();
    bool any = false;
    for (size_t k = 0; k < sims_.size(); ++k) {
        if (reinit == 1)
            initialised[k] = false;
        else if (reinit == 0)
            initialised[k] = true;
        if (initialised[k])
            init[k].q_count = 0;
        else
            any = true;
    }
    if (any)
        enqueue("clode_batch_initialize_observer", init);
    enqueue("clode_batch_features", args());
    std::fill(initialised.begin(), initialised.end(), true);
}

void Batch::begin() { hip_check(hipEventRecord(ev0_, stream_), "hipEventRecord"); }

void Batch::end()
{
    hip_check(hipEventRecord(ev1_, stream_), "hipEventRecord");
    hip_check(hipStreamSynchronize(stream_), "batch kernels");
    float ms = 0.f;
    hipEventElapsedTime(&ms, ev0_, ev1_);
    lastMs = ms;
}

void Batch::transient()
{
    prepare();
    begin();
    enqueue("clode_batch_transient", args());
    end();
}

void Batch::initializeObserver()
{
    if (!features_)
        throw InvalidArgument("initialize_observer needs feature simulators");
    prepare();
    begin();
    enqueue("clode_batch_initialize_observer", args());
    end();
    for (Simulator *s : sims_)
        s->observerInitialized = true;
}

void Batch::features(int reinit)
{
    if (!features_)
        throw InvalidArgument("features needs feature simulators");
    prepare();
    std::vector<bool> initialised;
    for (Simulator *s : sims_)
        initialised.push_back(s->observerInitialized);
    begin();
    enqueueFeatures(reinit, initialised);
    end();
    for (Simulator *s : sims_)
        s->observerInitialized = true;
}

void Batch::shiftX0()
{
    prepare();
    begin();
    enqueueShift(args());
    end();
}

void Batch::sweep(int rounds, bool graph)
{
    if (rounds < 1)
        throw InvalidArgument("sweep needs at least one round");
    if (!features_)
        throw InvalidArgument("sweep needs feature simulators");
    prepare();
    auto p = program();
    // load every kernel and query its residency before a capture begins
    for (const char *k : {"clode_queue_reset", "clode_batch_transient", "clode_batch_initialize_observer",
                          "clode_batch_features", "clode_batch_shift"})
        p->function(device_, k);
    for (const char *k : {"clode_batch_transient", "clode_batch_initialize_observer", "clode_batch_features"})
        gridFor(p->function(device_, k), 1);
    std::vector<bool> initialised;
    for (Simulator *s : sims_)
        initialised.push_back(s->observerInitialized);
    auto sequence = [&]() {
        for (int r = 0; r < rounds; ++r) {
            enqueue("clode_batch_transient", args());
            enqueueShift(args());
            enqueueFeatures(-1, initialised);
            enqueueShift(args());
        }
    };
    if (!graph) {
        begin();
        sequence();
        end();
    } else {
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        try {
            sequence();
        } catch (...) {
            hipStreamEndCapture(stream_, &g);
            if (g)
                hipGraphDestroy(g);
            throw;
        }
        hip_check(hipStreamEndCapture(stream_, &g), "hipStreamEndCapture");
        hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            hipGraphDestroy(g);
            hip_check(e, "hipGraphInstantiate");
        }
        try {
            begin();
            hip_check(hipGraphLaunch(exec, stream_), "hipGraphLaunch");
            end();
        } catch (...) {
            hipStreamSynchronize(stream_);
            hipGraphExecDestroy(exec);
            hipGraphDestroy(g);
            throw;
        }
        hipGraphExecDestroy(exec);
        hipGraphDestroy(g);
    }
    for (Simulator *s : sims_)
        s->observerInitialized = true;
}

// where the autotuner keeps its choice: the first simulator's code-object
// cache directory, keyed by the batched program, the device, the ensemble
// sizes, the stride and the step cap ("" when no directory is set)
std::string Batch::tunePath(int stride, unsigned int maxSteps)
{
    const std::string &dir = sims_[0]->root;
    if (dir.empty())
        return std::string();
    hipDeviceProp_t prop;
    std::string dev = hipGetDeviceProperties(&prop, device_) == hipSuccess ? prop.gcnArchName : "?";
    std::string key = progKey_ + "\n" + dev + "\n" + std::to_string(stride) + " " + std::to_string(maxSteps);
    for (Simulator *s : sims_)
        key += " " + std::to_string(s->nPts);
    uint64_t h = 1469598103934665603ull;
    for (unsigned char c : key) {
        h ^= c;
        h *= 1099511628211ull;
    }
    char name[64];
    std::snprintf(name, sizeof(name), "autotune_%016llx.txt", (unsigned long long)h);
    return dir + (dir.back() == '/' ? "" : "/") + name;
}

void Batch::autotune(int stride, unsigned int maxSteps, int &blockOut, int &sliceOut)
{
    if (stride < 1)
        throw InvalidArgument("autotune stride must be at least 1");
    prepare();
    program();
    tuneTable.clear();
    tuneFromCache = false;
    const std::string path = tunePath(stride, maxSteps);
    const char *fresh = std::getenv("CLODE_AUTOTUNE_FRESH");
    if (!path.empty() && !(fresh && std::strcmp(fresh, "1") == 0)) {
        std::ifstream f(path);
        int b = 0, sl = 0;
        if (f >> b >> sl && (b == 64 || b == 128 || b == 256) && sl >= 1) {
            block_ = blockOut = b;
            slice_ = sliceOut = sl;
            tuneFromCache = true;
            lastMs = 0.0;
            return;
        }
    }
    // per-member state the transient kernel overwrites
    struct Saved {
        DeviceBuffer xf, rng, dt, tf, steps;
    };
    std::vector<Saved> saved(sims_.size());
    auto copy = [&](void *dst, const void *src, size_t bytes) {
        if (bytes)
            hip_check(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream_), "hipMemcpy D2D");
    };
    for (size_t k = 0; k < sims_.size(); ++k) {
        Simulator *s = sims_[k];
        saved[k].xf.allocate(device_, s->d_xf.bytes);
        saved[k].rng.allocate(device_, s->d_rng.bytes);
        saved[k].dt.allocate(device_, s->d_dt.bytes);
        saved[k].tf.allocate(device_, s->d_tf.bytes);
        copy(saved[k].xf.ptr, s->d_xf.ptr, s->d_xf.bytes);
        copy(saved[k].rng.ptr, s->d_rng.ptr, s->d_rng.bytes);
        copy(saved[k].dt.ptr, s->d_dt.ptr, s->d_dt.bytes);
        copy(saved[k].tf.ptr, s->d_tf.ptr, s->d_tf.bytes);
        if (s->d_steps.ptr) {
            saved[k].steps.allocate(device_, s->d_steps.bytes);
            copy(saved[k].steps.ptr, s->d_steps.ptr, s->d_steps.bytes);
        }
    }
    auto restore = [&]() {
        for (size_t k = 0; k < sims_.size(); ++k) {
            Simulator *s = sims_[k];
            copy(s->d_xf.ptr, saved[k].xf.ptr, s->d_xf.bytes);
            copy(s->d_rng.ptr, saved[k].rng.ptr, s->d_rng.bytes);
            copy(s->d_dt.ptr, saved[k].dt.ptr, s->d_dt.bytes);
            copy(s->d_tf.ptr, saved[k].tf.ptr, s->d_tf.bytes);
            if (saved[k].steps.ptr)
                copy(s->d_steps.ptr, saved[k].steps.ptr, s->d_steps.bytes);
        }
    };
    const int blocks[] = {64, 128, 256};
    const int slices[] = {1, 4, 16, 64};
    const int reps = 3;
    std::vector<std::pair<int, int>> shapes;
    for (int b : blocks)
        for (int sl : slices)
            shapes.push_back({b, sl});
    std::vector<std::vector<double>> times(shapes.size());
    const int keepB = block_, keepS = slice_;
    bool warmed = false;
    try {
        // interleaved rounds: slow drifts of the clock or of the machine's
        // load touch every shape alike
        for (int rep = 0; rep < reps; ++rep)
            for (size_t c = 0; c < shapes.size(); ++c) {
                block_ = shapes[c].first;
                slice_ = shapes[c].second;
                std::vector<ClodeArgs> a = args();
                for (size_t k = 0; k < a.size(); ++k) {
                    a[k].q_stride = (unsigned int)stride;
                    a[k].q_count = (unsigned int)((sims_[k]->nPts + stride - 1) / stride);
                    if (maxSteps > 0)
                        a[k].sp_max_steps = std::min(a[k].sp_max_steps, maxSteps);
                }
                if (!warmed) {  // first launch loads the module and warms the caches
                    enqueue("clode_batch_transient", a);
                    restore();
                    warmed = true;
                }
                begin();
                enqueue("clode_batch_transient", a);
                end();
                restore();
                times[c].push_back(lastMs);
            }
    } catch (...) {
        hipStreamSynchronize(stream_);
        block_ = keepB;
        slice_ = keepS;
        throw;
    }
    hip_check(hipStreamSynchronize(stream_), "hipStreamSynchronize");
    size_t best = 0;
    for (size_t c = 0; c < shapes.size(); ++c) {
        std::vector<double> t = times[c];
        std::sort(t.begin(), t.end());
        tuneTable.push_back({shapes[c].first, shapes[c].second, t[t.size() / 2], t.front(), t.back()});
        if (tuneTable[c].median < tuneTable[best].median)
            best = c;
        log_msg(LOG_DEBUG, "autotune block " + std::to_string(shapes[c].first) + " slice " +
                               std::to_string(shapes[c].second) + ": median " + std::to_string(t[t.size() / 2]) +
                               " ms");
    }
    block_ = blockOut = shapes[best].first;
    slice_ = sliceOut = shapes[best].second;
    lastMs = tuneTable[best].median;
    if (!path.empty()) {
        mkdir(sims_[0]->root.c_str(), 0755);
        const std::string tmp = path + ".tmp" + std::to_string(getpid());
        bool ok = false;
        {
            std::ofstream f(tmp);
            ok = (bool)(f << block_ << " " << slice_ << "\n");
        }
        if (!ok || std::rename(tmp.c_str(), path.c_str()) != 0)
            std::remove(tmp.c_str());
    }
}

}  // namespace clode
