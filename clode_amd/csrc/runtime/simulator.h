// Simulator state and operations behind the C ABI.
//
// One class covers the three reference host classes (kind selects the
// extra members): CLODE (clode/cpp/CLODE.hpp:93-189), CLODEfeatures
// (clode/cpp/CLODEfeatures.hpp) and CLODEtrajectory
// (clode/cpp/CLODEtrajectory.hpp).  Host-side semantics are the reference's:
// size-mismatched inputs are ignored with an info message, the device step
// buffer is refilled only when the ensemble size changes, the RNG is reseeded
// (non-deterministically) whenever the ensemble size changes, observer data
// survive between features() calls until the ensemble size changes or a
// re-initialisation is requested, and every compute call blocks.
#pragma once

#include <map>

#include "clode_runtime.h"

namespace clode {

class Simulator {
public:
    Simulator(int kind, const ProblemInfo &pi, const std::string &stepper, const std::string &observer,
              const ObserverParams &op, int precision, int device, const std::string &root);
    ~Simulator();

    // configuration
    void setProblemInfo(const ProblemInfo &pi);
    void setStepper(const std::string &s);
    void setPrecision(int prec);
    void setDevice(int device);
    void build();
    void compileOnly();

    // data
    void setProblemData(const std::vector<double> &x0, const std::vector<double> &pars);
    void setTspan(const std::vector<double> &t);
    void setX0(const std::vector<double> &x0);
    void setPars(const std::vector<double> &pars);
    void setSolverParams(const SolverParams &sp);
    void seedRNG();
    void seedRNG(int seed);
    void setMemberTspan(const double *tspan_m, size_t n);

    // compute
    void transient();
    void shiftTspan();
    void shiftX0();

    // getters
    std::vector<double> getX0();
    std::vector<double> getXf();
    std::vector<double> getDt();
    std::vector<double> getTf();
    std::string programString();

    // features
    void setObserverParams(const ObserverParams &op);
    void setObserver(const std::string &name);
    void initializeObserver();
    void features();
    void features(bool reinit);
    void transientFeatures(double tr0, double tr1, double fe0, double fe1);
    std::vector<double> getF();
    // F into out (capacity cap): feature-major as getF, or member-major
    // ([nPts][nF], transposed on the device)
    void getFInto(double *out, size_t cap);
    void getFMembers(double *out, size_t cap);

    // trajectory
    void trajectory();
    std::vector<double> getT();
    std::vector<double> getX();
    std::vector<double> getDx();
    std::vector<double> getAux();
    std::vector<int> getNstored();

    // public state (read by the C ABI)
    int kind;
    ProblemInfo prob;
    std::string rhsSource;
    std::string stepper;
    int prec;
    int device;
    std::string root;
    SolverParams sp;
    std::vector<double> tspan{0.0, 1000.0};
    int nPts = 0;
    std::vector<double> x0, pars, dtHost;
    std::vector<unsigned long long> rngHost;

    std::string observer;
    ObserverParams op;
    bool observerInitialized = false;
    int nFeatures = 0;
    std::vector<std::string> featureNames;

    unsigned int nStoreMax = 0;
    size_t telements = 0, xelements = 0, auxelements = 0;

    int launchMode = 0;   // 0 grid, 1 persistent work queue
    int blockSize = 0;    // 0 = auto
    unsigned int queueSlice = 8;  // work queue: loop iterations between queue checks
    bool dtRefresh = false;
    double lastKernelMs = 0.0;
    // waves per SIMD the last launch's kernel was compiled for (0: the
    // default compilation)
    int lastVariantWaves = 0;
    bool stepCounter = false;

    // Features replay, off unless a budget is set: with a two-pass observer
    // (thresh2, nhood2) and no noise, the warm-up pass also records every
    // member's accepted steps, up to `bytes` of device memory, and the
    // features pass that follows it with nothing changed in between reads
    // them back instead of integrating them again (bit-identical results).
    void setReplayBudget(size_t bytes);
    size_t replayBudget = 0;
    unsigned int replayCap = 0;  // records per member of the last recording (0: none)
    bool lastReplay = false;     // the last features pass replayed a record

    void setStepCounter(bool on);
    void resetStepCounts();
    // row 0: accepted steps per member; row 1: step trials (a step and its rejected trials)
    std::vector<unsigned long long> getStepCounts(int row = 0);
    std::vector<unsigned long long> getRngState();

private:
    friend class Batch;
    DeviceBuffer d_x0, d_pars, d_xf, d_rng, d_dt, d_tf;
    DeviceBuffer d_odr, d_odu, d_F;
    DeviceBuffer d_t, d_x, d_dx, d_aux, d_nStored;
    DeviceBuffer d_work, d_tspan_m, d_steps, d_qorder;
    DeviceBuffer d_replay, d_replay_n;
    // inputGen counts the calls that change what an integration reads; a
    // record is replayed only while it still equals replayGen
    unsigned long long inputGen = 0, replayGen = ~0ull;
    bool replayUsable() const;
    bool prepareReplay();
    ClodeArgs replayArgs();
    size_t odrFields = 0, oduFields = 0, Felements = 0;
    bool memberTspan = false;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::shared_ptr<Program> program;
    std::string programKeyCache;

    size_t realSize() const { return prec == PREC_DOUBLE ? 8 : 4; }   // pars, trajectory x
    size_t wideSize() const { return prec == PREC_SINGLE ? 4 : 8; }   // x0/xf/dt/tf/F/odata/traj t
    int observerId() const;
    ProgramConfig programConfig() const;
    std::shared_ptr<Program> currentProgram();
    // the same program compiled for at least `waves` waves per SIMD (0: the
    // default compilation), used by grid launches of large ensembles (see
    // gridFunction), and with the replay kernels when `replay` is set
    std::shared_ptr<Program> variantProgram(int waves, bool replay);
    std::map<int, std::pair<std::string, std::shared_ptr<Program>>> programsW;
    hipFunction_t gridFunction(const char *kernel, unsigned int nThreads, unsigned int block, bool replay = false);
    void ensureStream();
    void setNpts(int n);
    void upload(DeviceBuffer &b, const std::vector<double> &v, size_t elem);
    std::vector<double> download(const DeviceBuffer &b, size_t n, size_t elem);
    // n elements of b (elem bytes each) as double into out
    void downloadInto(const DeviceBuffer &b, size_t n, size_t elem, double *out);
    DeviceBuffer d_Fm;  // member-major copy of F for getFMembers
    ClodeArgs makeArgs();
    void launch(const char *kernel, const ClodeArgs &a, unsigned int nThreads, bool replay = false);
    void updateObserverInfo();
    void resizeFeaturesVariables();
    void resizeTrajectoryVariables();
    void requireData(const char *what) const;
};

// resident-workgroup grid for a persistent work-queue launch of fn
unsigned int queue_grid(hipFunction_t fn, int device, unsigned int block, size_t members);

}  // namespace clode
