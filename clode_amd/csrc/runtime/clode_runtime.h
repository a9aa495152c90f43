// Internal C++ interface of the HIP runtime behind include/clode_hip.h.
//
// Replaces the reference's OpenCL host core: CLODE (clode/cpp/CLODE.cpp),
// CLODEfeatures (clode/cpp/CLODEfeatures.cpp), CLODEtrajectory
// (clode/cpp/CLODEtrajectory.cpp) and OpenCLResource
// (clode/cpp/OpenCLResource.cpp).  Programs are compiled with hipRTC for
// gfx950 from the embedded device sources plus the user's right-hand side and
// cached by content hash (in memory and on disk).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../device/clode_kernel_args.h"
#include "../device/clode_layout.h"

namespace clode {

// ------------------------------------------------------------------ logging
enum LogLevel { LOG_TRACE = 0, LOG_DEBUG, LOG_INFO, LOG_WARN, LOG_ERR, LOG_CRITICAL, LOG_OFF };
void log_msg(int level, const std::string &msg);
void set_log_level(int level);
int get_log_level();

// ------------------------------------------------------------------- errors
struct InvalidArgument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};
struct RuntimeError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
void hip_check(hipError_t e, const char *what);

// ------------------------------------------------------------ data structs
struct ProblemInfo {
    std::string src_file;
    int nVar = 0, nPar = 0, nAux = 0, nWiener = 0;
    std::vector<std::string> varNames, parNames, auxNames;
};

struct SolverParams {
    double dt = 0.1, dtmax = 0.5, abstol = 1e-6, reltol = 1e-3;
    unsigned int max_steps = 1000000, max_store = 1000000, nout = 1;
};

struct ObserverParams {
    unsigned int eVarIx = 0, fVarIx = 0, maxEventCount = 100, maxEventTimestamps = 0;
    double minXamp = 0., minIMI = 0., nHoodRadius = 0.05, xUpThresh = 0.2, xDownThresh = 0.2;
    double dxUpThresh = 0., dxDownThresh = 0., eps_dx = 0.;
};

enum Precision { PREC_SINGLE = 1, PREC_DOUBLE = 2, PREC_FP32C = 3 };

// registries (clode/cpp/steppers.cl:26-43, clode/cpp/observers.cl:122-143)
const std::map<std::string, int> &stepper_map();
const std::map<std::string, int> &observer_map();
std::vector<std::string> observer_feature_names(int obs, const ProblemInfo &pi, int nStore, unsigned int fVar);

// ------------------------------------------------------------------- device
struct DeviceBuffer {
    void *ptr = nullptr;
    size_t bytes = 0;
    int device = -1;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer &) = delete;
    DeviceBuffer &operator=(const DeviceBuffer &) = delete;
    ~DeviceBuffer() { release(); }
    void allocate(int dev, size_t nbytes);
    void release();
    template <class T>
    T *as() const { return static_cast<T *>(ptr); }
};

void set_device(int device);
int device_count();

// ------------------------------------------------------------------ program
struct ProgramConfig {
    int prec = PREC_DOUBLE;
    int stepper = 3;
    int observer = 0;  // 0 = none
    int eVar = 0, fVar = 0;  // observer event / feature variable
    int nVar = 0, nPar = 0, nAux = 0, nWiener = 0, nStore = 0;
    int wavesPerEU = 0;  // > 0: kernels compiled for at least this many waves per SIMD
    bool replay = false;  // also the warm-up record / features replay kernels
    std::string rhs;
    std::string rhs_name;
};

class Program {
public:
    std::string source;
    std::vector<std::string> options;
    std::vector<char> code;
    hipFunction_t function(int device, const char *name);
    ~Program();

private:
    std::mutex mu_;
    std::map<int, hipModule_t> modules_;
};

std::string compose_source(const ProgramConfig &cfg);
// toolchain text mixed into every program's cache key
std::string toolchain_identity();
std::vector<std::string> compile_options(const ProgramConfig &cfg);
std::shared_ptr<Program> get_program(const ProgramConfig &cfg, const std::string &cache_dir);
std::shared_ptr<Program> get_program_source(const std::string &source, const std::vector<std::string> &options,
                                            const std::string &cache_dir);

// batched programs: up to CLODE_BATCH_MAX systems in one code object
constexpr size_t CLODE_BATCH_MAX = 8;
std::string compose_batch_source(const std::vector<ProgramConfig> &cfgs);

}  // namespace clode
