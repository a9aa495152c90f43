// Kernel argument block shared by the host runtime (g++/clang++, x86-64) and
// the hipRTC device programs (gfx950).  Passed by value as the single kernel
// argument; every field has the same size and alignment on both targets.
// Scalars are carried in double and rounded to the program's working
// precision inside the kernel, which is the same round-to-nearest conversion
// the reference applies on the host before copying (clode/cpp/CLODE.cpp:405-417,
// CLODEfeatures.cpp:125-141).
#ifndef CLODE_KERNEL_ARGS_H
#define CLODE_KERNEL_ARGS_H

struct ClodeArgs {
    // time span (clode/cpp/transient.cl:10)
    double t0;
    double t1;
    // SolverParams (clode/cpp/clODE_struct_defs.cl:11-20)
    double sp_dt;
    double sp_dtmax;
    double sp_abstol;
    double sp_reltol;
    unsigned int sp_max_steps;
    unsigned int sp_max_store;
    unsigned int sp_nout;
    unsigned int nPts;
    // ObserverParams (clode/cpp/observers.cl:25-46)
    unsigned int op_eVarIx;
    unsigned int op_fVarIx;
    unsigned int op_maxEventCount;
    unsigned int op_maxEventTimestamps;
    double op_minXamp;
    double op_minIMI;
    double op_nHoodRadius;
    double op_xUpThresh;
    double op_xDownThresh;
    double op_dxUpThresh;
    double op_dxDownThresh;
    double op_eps_dx;
    // member arrays, structure of arrays: element [j][i] at j*nPts + i
    void *x0;            // [nVar][nPts]   state in  (x store type)
    void *pars;          // [nPar][nPts]   realtype
    void *xf;            // [nVar][nPts]   state out (x store type)
    unsigned long long *rng;  // [2][nPts]  xoroshiro128+ words
    void *dt;            // [nPts]          step size in/out
    void *tf;            // [nPts]          final time
    void *odr;           // [nReal][nPts]   observer reals
    unsigned int *odu;   // [nUint][nPts]   observer unsigned ints
    void *F;             // [nFeat][nPts]   features
    void *tr_t;          // [maxStore][nPts]
    void *tr_x;          // [maxStore][nVar][nPts]
    void *tr_dx;         // [maxStore][nVar][nPts]
    void *tr_aux;        // [maxStore][nAux][nPts]
    int *nStored;        // [nPts]
    void *tspan_m;               // optional per-member [2][nPts] time spans (double)
    unsigned long long *steps;   // optional [2][nPts]: accepted steps, then step trials (accumulated)
    // work queue (persistent launches): position q -> member q_order[q*q_stride]
    unsigned int *work_counter;  // next queue position to claim; zeroed before each launch
    const unsigned int *q_order; // optional queue order (member indices)
    unsigned int q_count;        // queue positions
    unsigned int q_stride;       // stride through the order (1 = every member)
    unsigned int q_slice;        // loop iterations between queue checks
    unsigned int q_pad;
    // warm-up record (replay kernels only): the warm-up pass writes up to
    // replay_cap records per member and their count, the features pass reads
    // them back instead of integrating those steps again
    double *replay;              // [replay_cap / CHUNK][nPts][CHUNK][record]
    unsigned int *replay_n;      // [nPts] records written
    unsigned int replay_cap;     // records per member (a multiple of the chunk)
    unsigned int replay_pad;
};

// Arguments of the one-launch transient -> features sequence (clode_sequence):
// the same buffers, the transient's span in tr and the features' span in fe.
struct ClodeSeqArgs {
    ClodeArgs tr;
    ClodeArgs fe;
};

#ifdef __cplusplus
static_assert(sizeof(void *) == 8, "64-bit targets only");
#endif

#endif  // CLODE_KERNEL_ARGS_H
