// Layout of the per-member observer state and feature vectors.
//
// Shared by the host runtime (buffer sizing, feature names) and by the device
// programs (SoA addressing).  Observer ids for the names registered in
// clode/cpp/observers.cl:122-143: thresh2=1, localmax=2, nhood2=3, basic=4,
// basicall=5, nhood1=6.
//
// Observer state is stored structure-of-arrays: real field f of member i is
// at odr[f*nPts + i], unsigned field g at odu[g*nPts + i].  The first
// clode_obs_nscalar() real fields are the observer's scalars and per-variable
// arrays; they are followed by the event-time lists (clode_obs_nlists() lists
// of nStore entries each), which live only in HBM and are written when an event
// occurs.
#ifndef CLODE_LAYOUT_H
#define CLODE_LAYOUT_H

#define CLODE_OBS_THRESH2 1
#define CLODE_OBS_LOCALMAX 2
#define CLODE_OBS_NHOOD2 3
#define CLODE_OBS_BASIC 4
#define CLODE_OBS_BASICALL 5
#define CLODE_OBS_NHOOD1 6

// real scalar fields (per-variable arrays included; 3-point buffers only for
// the feature / event variable)
static inline constexpr int clode_obs_nscalar(int obs, int nVar, int nAux)
{
    return obs == CLODE_OBS_THRESH2    ? 3 + 3 + 2 + 5 * nVar + 3 * nAux + 21 + 1 + 8 + 6
           : obs == CLODE_OBS_LOCALMAX ? 3 + 3 + 2 + 5 * nVar + 3 * nAux + 6 + 4
           : obs == CLODE_OBS_NHOOD2   ? 3 + 2 + 2 + 7 * nVar + 3 * nAux + 9 + 3
           : obs == CLODE_OBS_BASIC    ? 5 + 2
           : obs == CLODE_OBS_BASICALL ? 5 * nVar + 3 * nAux + 2
           : obs == CLODE_OBS_NHOOD1   ? 3 + 3 * nVar + 4 + nVar + 5 * nVar + 3 * nAux + 9 + 2
                                       : 0;
}

static inline constexpr int clode_obs_nlists(int obs)
{
    return obs == CLODE_OBS_THRESH2 ? 2 : obs == CLODE_OBS_LOCALMAX ? 4 : obs == CLODE_OBS_NHOOD2 ? 1 : 0;
}

static inline constexpr int clode_obs_nreal(int obs, int nVar, int nAux, int nStore)
{
    return clode_obs_nscalar(obs, nVar, nAux) + clode_obs_nlists(obs) * nStore;
}

static inline constexpr int clode_obs_nuint(int obs)
{
    return obs == CLODE_OBS_THRESH2    ? 4
           : obs == CLODE_OBS_LOCALMAX ? 2
           : obs == CLODE_OBS_NHOOD2   ? 5
           : obs == CLODE_OBS_BASIC    ? 1
           : obs == CLODE_OBS_BASICALL ? 1
           : obs == CLODE_OBS_NHOOD1   ? 5
                                       : 0;
}

// feature counts (clode/cpp/observers/observer_threshold_2.clh:15-53,
// observer_local_maximum.clh:19-45, observer_neighborhood_2.clh:14-41,
// observer_basic.clh:6-14, observer_basic_allVar.clh:6-27,
// observer_neighborhood_1.clh:6-37)
static inline constexpr int clode_obs_nfeatures(int obs, int nVar, int nAux, int nStore)
{
    return obs == CLODE_OBS_THRESH2    ? 18 + 5 * nVar + 3 * nAux + 2 * nStore + 5
           : obs == CLODE_OBS_LOCALMAX ? 6 + 5 * nVar + 3 * nAux + 4 * nStore + 2
           : obs == CLODE_OBS_NHOOD2   ? 6 + 7 * nVar + 3 * nAux + nStore + 5
           : obs == CLODE_OBS_BASIC    ? 6
           : obs == CLODE_OBS_BASICALL ? 5 * nVar + 3 * nAux + 1
           : obs == CLODE_OBS_NHOOD1   ? 6 + 5 * nVar + 3 * nAux + 5
                                       : 0;
}

static inline constexpr bool clode_obs_two_pass(int obs)
{
    return obs == CLODE_OBS_THRESH2 || obs == CLODE_OBS_NHOOD2;
}

// Warm-up record for the features replay of a two-pass observer: for every
// accepted step of the warm-up pass, the member's time, state, derivative,
// aux values and next step size, and in fp32c (prec 3) the state's
// compensation term, each as a double (every precision's values widen to
// double exactly).  A member's records are grouped in chunks of
// CLODE_REPLAY_CHUNK consecutive steps, and chunk c of member i is at
// (c * nPts + i) * CLODE_REPLAY_CHUNK * record doubles: the lanes of a wave,
// which replay their steps in lockstep, read neighbouring chunks.
#define CLODE_REPLAY_CHUNK 8
static inline constexpr int clode_replay_record(int nVar, int nAux, int prec)
{
    return 2 + 2 * nVar + nAux + (prec == 3 ? nVar : 0);
}

#endif  // CLODE_LAYOUT_H
