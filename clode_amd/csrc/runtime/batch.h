// Batched ensembles: several simulators, possibly with different right-hand
// sides, dimensions, steppers and observers, integrated by one persistent
// work-queue launch.
//
// The reference runs one system per program and one time span per launch
// (clode/cpp/CLODE.cpp:148, transient.cl:10); a batch has no counterpart
// there.  Each member simulator keeps its own buffers, parameters, noise
// stream and host semantics; the batch compiles all of their systems into one
// code object (one namespace per system) and launches dispatching kernels
// whose lanes claim members from per-system work queues.  Per-member time
// spans (Simulator::setMemberTspan) are honoured, with the queue ordered
// longest span first.  A member's results are bit-identical to running its
// simulator on its own, because the same lane code integrates it.
#pragma once

#include "simulator.h"

namespace clode {

class Batch {
public:
    explicit Batch(const std::vector<Simulator *> &sims);
    ~Batch();

    // workgroup size (64..256, multiple of 64) and loop iterations between
    // queue checks (>= 1)
    void setLaunch(int block, int slice);
    int block() const { return block_; }
    int slice() const { return slice_; }

    void transient();
    void initializeObserver();
    // reinit: -1 continue (initialise where needed), 0 continue, 1 re-initialise
    // (the semantics of features() / features(false) / features(true))
    void features(int reinit);
    void shiftX0();

    // rounds x (transient, x0 <- xf, features, x0 <- xf); with graph = true the
    // whole sequence is captured into one graph and launched once
    void sweep(int rounds, bool graph);

    // time the transient work-queue kernel on every stride-th queued member,
    // each trial capped at max_steps steps per member (0: no cap), for each
    // candidate launch shape, three times in interleaved rounds, and keep the
    // shape with the fastest median (member state restored afterwards);
    // returns the chosen shape.  The choice is kept in the code-object cache
    // directory (keyed by program, device, ensemble sizes, stride and cap) and
    // reused by later runs without trials (CLODE_AUTOTUNE_FRESH=1 re-times).
    void autotune(int stride, unsigned int maxSteps, int &block, int &slice);
    // the last autotune's trials: block, slice, median, min, max (ms); empty
    // when the choice came from the cache
    struct TuneRow {
        int block, slice;
        double median, min, max;
    };
    std::vector<TuneRow> tuneTable;
    bool tuneFromCache = false;

    // compile the batched program for these simulators (no device needed)
    static void compileOnly(const std::vector<Simulator *> &sims);

    double lastMs = 0.0;   // device time of the last operation
    size_t size() const { return sims_.size(); }

private:
    std::vector<Simulator *> sims_;
    int device_ = 0, prec_ = 0;
    bool features_ = false;
    int block_ = 64, slice_ = 8;
    std::shared_ptr<Program> prog_;
    std::string progKey_;
    hipStream_t stream_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    DeviceBuffer counters_;
    std::map<std::pair<void *, int>, size_t> resident_;

    std::shared_ptr<Program> program();
    unsigned int gridFor(hipFunction_t fn, size_t members);
    void prepare();
    std::vector<ClodeArgs> args();
    void enqueue(const char *kernel, std::vector<ClodeArgs> a);
    void enqueueShift(std::vector<ClodeArgs> a);
    void enqueueFeatures(int reinit, std::vector<bool> &initialised);
    void begin();
    void end();
    std::string tunePath(int stride, unsigned int maxSteps);
};

}  // namespace clode
