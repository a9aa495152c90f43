// On-the-fly feature extraction ("observers"), one member per lane.
//
// Protocol (clode/cpp/observers.cl:4-14): init, warmup (two-pass observers),
// initEventDetector, update (every accepted step), eventFunction,
// computeEventFeatures (returns true at the terminal event count),
// finalizeFeatures, finalizeObserverData (time shift for continuation).
//
//   thresh2  clode/cpp/observers/observer_threshold_2.clh:121-470
//   localmax clode/cpp/observers/observer_local_maximum.clh:92-328
//   nhood2   clode/cpp/observers/observer_neighborhood_2.clh:90-313
//   basic    clode/cpp/observers/observer_basic.clh:17-90
//   basicall clode/cpp/observers/observer_basic_allVar.clh:30-124
//   nhood1   clode/cpp/observers/observer_neighborhood_1.clh:40-310
//   helpers  clode/cpp/clODE_utilities.cl:31-177 (runningMean,
//            runningMeanTime, first-occurrence arg-max/arg-min, norm_2)
//
// Differences from the reference's layout (not its arithmetic):
//  * the observer struct lives in registers during the step loop and is
//    spilled structure-of-arrays between launches (see clode_layout.h);
//  * of the reference's 3-point x/dx buffers only those of the feature
//    variable (and, for nhood2, the event variable) are kept: the others are
//    never read;
//  * event-time lists (N_STORE_EVENTS entries) are written straight to HBM
//    when the event happens instead of being carried in registers;
//  * thresh2's local-extremum scratch buffer has 3 entries (the reference
//    declares N_VAR entries and writes 3, out of bounds for N_VAR < 3);
//  * localmax skips the local-minimum list store while eventcount == 0 (the
//    reference writes element eventcount-1 = UINT_MAX there).
// eVarIx / fVarIx are compile-time constants of the program (CLODE_EVAR,
// CLODE_FVAR; the host keys its program cache on them), so the selected
// variables are plain registers and no scratch memory is ever indexed.
//
// fp32-compensated mode (CLODE_PREC 3): time fields and event statistics are
// double, per-step time means are kept as double integrals S = sum x dt and
// divided once at finalisation (mean = S / elapsed), the mean step is
// elapsed/steps, removing the per-step divisions of the running means.
#ifndef CLODE_OBSERVERS_H
#define CLODE_OBSERVERS_H

#include "clode_layout.h"

#ifndef N_STORE_EVENTS
#define N_STORE_EVENTS 0
#endif

#if CLODE_PREC == 3
typedef double obsreal;  // statistics and times
typedef double ostore;   // element type of the HBM observer arrays and F
#else
typedef realtype obsreal;
typedef realtype ostore;
#endif

// Observer parameters in working precision (clode/cpp/observers.cl:25-46)
struct ObserverParamsR {
    unsigned int eVarIx;
    unsigned int fVarIx;
    unsigned int maxEventCount;
    unsigned int maxEventTimestamps;
    realtype minXamp;
    realtype minIMI;
    realtype nHoodRadius;
    realtype xUpThresh;
    realtype xDownThresh;
    realtype dxUpThresh;
    realtype dxDownThresh;
    realtype eps_dx;
};

#ifndef CLODE_EVAR
#define CLODE_EVAR 0
#endif
#ifndef CLODE_FVAR
#define CLODE_FVAR 0
#endif

CLODE_DEV void runningMean(obsreal &mean, obsreal newValue, unsigned int eventCount)
{
    if (eventCount == 1)
        mean = newValue;
    else if (eventCount > 1)
        mean += (newValue - mean) / (obsreal)eventCount;
}

CLODE_DEV realtype runningMeanTime(realtype mean, realtype newValue, realtype dt, realtype total_delta)
{
    return mean + (newValue - mean) * dt / total_delta;
}

// HBM access for one member's persistent observer state
struct ObsIO {
    ostore *r;
    unsigned int *u;
    unsigned int nPts;
    unsigned int i;
    unsigned int ri;
    unsigned int ui;
};

struct ObsLoad {
    ObsIO io;
    template <class T>
    CLODE_DEV void real(T &v) { v = (T)io.r[(size_t)(io.ri++) * io.nPts + io.i]; }
    CLODE_DEV void uint(unsigned int &v) { v = io.u[(size_t)(io.ui++) * io.nPts + io.i]; }
};

struct ObsStore {
    ObsIO io;
    template <class T>
    CLODE_DEV void real(T &v) { io.r[(size_t)(io.ri++) * io.nPts + io.i] = (ostore)v; }
    CLODE_DEV void uint(unsigned int &v) { io.u[(size_t)(io.ui++) * io.nPts + io.i] = v; }
};

// event-time list element: list L, entry j
CLODE_DEV void obs_list_store(const ObsIO &io, int nscalar, int L, unsigned int j, obsreal v)
{
    io.r[(size_t)(nscalar + L * N_STORE_EVENTS + (int)j) * io.nPts + io.i] = (ostore)v;
}
CLODE_DEV ostore obs_list_load(const ObsIO &io, int nscalar, int L, int j)
{
    return io.r[(size_t)(nscalar + L * N_STORE_EVENTS + j) * io.nPts + io.i];
}

// feature output F[ix*nPts + i]
struct FeatOut {
    ostore *F;
    unsigned int nPts;
    unsigned int i;
    int ix;
    CLODE_DEV void put(obsreal v) { F[(size_t)(ix++) * nPts + i] = (ostore)v; }
};

// ------------------------------------------------------------------------
#if CLODE_OBSERVER == CLODE_OBS_THRESH2
#define CLODE_TWO_PASS 1
struct Observer {
    static constexpr int NSCALAR = clode_obs_nscalar(CLODE_OBS_THRESH2, N_VAR, N_AUX);
    timetype tb0, tb1, tb2;
    realtype xb0, xb1, xb2;
    realtype dxb1, dxb2;
    realtype xMax[N_VAR], xMin[N_VAR], dxMax[N_VAR], dxMin[N_VAR];
    obsreal xMean[N_VAR];  // running mean (plain) or integral (fp32c)
    realtype auxMax[N_AUX1], auxMin[N_AUX1];
    obsreal auxMean[N_AUX1];
    obsreal nMaxima[3], period[3], upDuration[3], downDuration[3], duty[3], activeDip[3], stepDt[3];
    obsreal fVarDownstateMean;  // running mean (plain) or integral since the down transition (fp32c)
    realtype xGlobalMax, xGlobalMin, dxGlobalMax, dxGlobalMin, xUp, xDown, dxUp, dxDown;
    timetype t_start, tLastEvent, tThisDown, tLastMax, tLastMin;
    realtype xLastMin;
    unsigned int thisNMaxima, stepcount, eventcount, inUpstate;

    template <class F>
    CLODE_DEV void fields(F &f)
    {
        f.real(tb0); f.real(tb1); f.real(tb2);
        f.real(xb0); f.real(xb1); f.real(xb2);
        f.real(dxb1); f.real(dxb2);
#pragma unroll
        for (int j = 0; j < N_VAR; j++) {
            f.real(xMax[j]); f.real(xMin[j]); f.real(xMean[j]); f.real(dxMax[j]); f.real(dxMin[j]);
        }
#pragma unroll
        for (int j = 0; j < N_AUX; j++) {
            f.real(auxMax[j]); f.real(auxMin[j]); f.real(auxMean[j]);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            f.real(nMaxima[j]); f.real(period[j]); f.real(upDuration[j]); f.real(downDuration[j]);
            f.real(duty[j]); f.real(activeDip[j]); f.real(stepDt[j]);
        }
        f.real(fVarDownstateMean);
        f.real(xGlobalMax); f.real(xGlobalMin); f.real(dxGlobalMax); f.real(dxGlobalMin);
        f.real(xUp); f.real(xDown); f.real(dxUp); f.real(dxDown);
        f.real(t_start); f.real(tLastEvent); f.real(tThisDown); f.real(tLastMax); f.real(tLastMin);
        f.real(xLastMin);
        f.uint(thisNMaxima); f.uint(stepcount); f.uint(eventcount); f.uint(inUpstate);
    }

    CLODE_DEV void init(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                        const ObserverParamsR &op, const ObsIO &io)
    {
        tb2 = ti;
        xb2 = xi[CLODE_FVAR];
        dxb2 = dxi[CLODE_FVAR];
        tb0 = tb1 = ti;
        xb0 = xb1 = xb2;
        dxb1 = dxb2;
        for (int j = 0; j < N_VAR; ++j) {
            xMean[j] = ZERO;
            xMax[j] = -BIG_REAL;
            xMin[j] = BIG_REAL;
            dxMax[j] = -BIG_REAL;
            dxMin[j] = BIG_REAL;
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = -BIG_REAL;
            auxMin[j] = BIG_REAL;
            auxMean[j] = ZERO;
        }
        for (int j = 0; j < N_STORE_EVENTS; ++j) {
            obs_list_store(io, NSCALAR, 0, j, ZERO);
            obs_list_store(io, NSCALAR, 1, j, ZERO);
        }
        nMaxima[0] = -BIG_REAL; nMaxima[1] = BIG_REAL; nMaxima[2] = ZERO;
        period[0] = -BIG_REAL; period[1] = BIG_REAL; period[2] = ZERO;
        upDuration[0] = -BIG_REAL; upDuration[1] = BIG_REAL; upDuration[2] = ZERO;
        downDuration[0] = -BIG_REAL; downDuration[1] = BIG_REAL; downDuration[2] = ZERO;
        duty[0] = -BIG_REAL; duty[1] = BIG_REAL; duty[2] = ZERO;
        activeDip[0] = -BIG_REAL; activeDip[1] = BIG_REAL; activeDip[2] = ZERO;
        stepDt[0] = -BIG_REAL; stepDt[1] = BIG_REAL; stepDt[2] = ZERO;
        fVarDownstateMean = ZERO;
        xGlobalMax = -BIG_REAL;
        xGlobalMin = BIG_REAL;
        dxGlobalMax = -BIG_REAL;
        dxGlobalMin = BIG_REAL;
        xUp = ZERO;
        xDown = ZERO;
        dxUp = ZERO;
        dxDown = ZERO;
        t_start = ti;
        tLastEvent = ZERO;
        tThisDown = ZERO;
        tLastMax = ZERO;
        tLastMin = ZERO;
        xLastMin = BIG_REAL;
        inUpstate = 0;
        thisNMaxima = 0;
        stepcount = 0;
        eventcount = 0;
    }

    CLODE_DEV void warmup(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                          const ObserverParamsR &op)
    {
        realtype xe = xi[CLODE_EVAR];
        realtype dxe = dxi[CLODE_EVAR];
        xGlobalMax = fmax(xGlobalMax, xe);
        xGlobalMin = fmin(xGlobalMin, xe);
        dxGlobalMax = fmax(dxGlobalMax, dxe);
        dxGlobalMin = fmin(dxGlobalMin, dxe);
    }

    CLODE_DEV void initEventDetector(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                                     const ObserverParamsR &op)
    {
        realtype xTrajectoryAmp = xGlobalMax - xGlobalMin;
        xUp = xGlobalMin + op.xUpThresh * xTrajectoryAmp;
        if (op.xDownThresh > ZERO)
            xDown = xGlobalMin + op.xDownThresh * xTrajectoryAmp;
        else
            xDown = xUp;
        dxUp = op.dxUpThresh * dxGlobalMax;
        if (op.dxDownThresh > ZERO)
            dxDown = op.dxDownThresh * dxGlobalMin;
        else
            dxDown = dxGlobalMin;
        inUpstate = xi[CLODE_EVAR] > xUp ? 1 : 0;
    }

    CLODE_DEV bool eventFunction(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                                 const ObserverParamsR &op)
    {
        if (stepcount < 2)
            return false;
        if (xGlobalMax - xGlobalMin < op.minXamp)
            return false;
        if (inUpstate)
            return false;
        return (xi[CLODE_EVAR] > xUp && dxi[CLODE_EVAR] > dxUp);
    }

    CLODE_DEV obsreal downMean(timetype ti, const realtype xi[], const ObserverParamsR &op)
    {
#if CLODE_PREC == 3
        double el = ti - tThisDown;
        return el > 0.0 ? fVarDownstateMean / el : (double)xi[CLODE_FVAR];
#else
        return fVarDownstateMean;
#endif
    }

    CLODE_DEV bool computeEventFeatures(timetype ti, const realtype xi[], const realtype dxi[],
                                        const realtype auxi[], const ObserverParamsR &op, const ObsIO &io)
    {
        timetype tThisEvent;
        ++eventcount;
        inUpstate = 1;
        tThisEvent = ti;
        if (eventcount > 1) {
            int nPeriods = eventcount - 1;
            nMaxima[0] = fmax((obsreal)thisNMaxima, nMaxima[0]);
            nMaxima[1] = fmin((obsreal)thisNMaxima, nMaxima[1]);
            runningMean(nMaxima[2], (obsreal)thisNMaxima, nPeriods);
            obsreal thisPeriod = tThisEvent - tLastEvent;
            period[0] = fmax(thisPeriod, period[0]);
            period[1] = fmin(thisPeriod, period[1]);
            runningMean(period[2], thisPeriod, nPeriods);
            obsreal thisUpDuration = tThisDown - tLastEvent;
            upDuration[0] = fmax(thisUpDuration, upDuration[0]);
            upDuration[1] = fmin(thisUpDuration, upDuration[1]);
            runningMean(upDuration[2], thisUpDuration, nPeriods);
            obsreal thisDownDuration = tThisEvent - tThisDown;
            downDuration[0] = fmax(thisDownDuration, downDuration[0]);
            downDuration[1] = fmin(thisDownDuration, downDuration[1]);
            runningMean(downDuration[2], thisDownDuration, nPeriods);
            obsreal thisduty = thisUpDuration / thisPeriod;
            duty[0] = fmax(thisduty, duty[0]);
            duty[1] = fmin(thisduty, duty[1]);
            runningMean(duty[2], thisduty, nPeriods);
            obsreal thisActiveDip = downMean(ti, xi, op) - xLastMin;
            activeDip[0] = fmax(thisActiveDip, activeDip[0]);
            activeDip[1] = fmin(thisActiveDip, activeDip[1]);
            runningMean(activeDip[2], thisActiveDip, nPeriods);
        }
        if (eventcount <= N_STORE_EVENTS)
            obs_list_store(io, NSCALAR, 0, eventcount - 1, tThisEvent);
        tLastEvent = tThisEvent;
        thisNMaxima = 0;
        return eventcount == op.maxEventCount;
    }

    CLODE_DEV void update(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                          const ObserverParamsR &op, const ObsIO &io)
    {
        ++stepcount;
        tb0 = tb1;
        tb1 = tb2;
        tb2 = ti;
        xb0 = xb1;
        xb1 = xb2;
        xb2 = xi[CLODE_FVAR];
        dxb1 = dxb2;
        dxb2 = dxi[CLODE_FVAR];
#if CLODE_PREC == 3
        double dt = tb2 - tb1;
        stepDt[0] = fmax(dt, stepDt[0]);
        stepDt[1] = fmin(dt, stepDt[1]);
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmaxf(xi[j], xMax[j]);
            xMin[j] = fminf(xi[j], xMin[j]);
            xMean[j] = fma((double)xi[j], dt, xMean[j]);
            dxMax[j] = fmaxf(dxi[j], dxMax[j]);
            dxMin[j] = fminf(dxi[j], dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = fmaxf(auxi[j], auxMax[j]);
            auxMin[j] = fminf(auxi[j], auxMin[j]);
            auxMean[j] = fma((double)auxi[j], dt, auxMean[j]);
        }
#else
        realtype dt = tb2 - tb1;
        realtype elapsedTime = ti - t_start;
        stepDt[0] = fmax(dt, stepDt[0]);
        stepDt[1] = fmin(dt, stepDt[1]);
        runningMean(stepDt[2], dt, stepcount);
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmax(xi[j], xMax[j]);
            xMin[j] = fmin(xi[j], xMin[j]);
            xMean[j] = runningMeanTime(xMean[j], xi[j], dt, elapsedTime);
            dxMax[j] = fmax(dxi[j], dxMax[j]);
            dxMin[j] = fmin(dxi[j], dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = fmax(auxi[j], auxMax[j]);
            auxMin[j] = fmin(auxi[j], auxMin[j]);
            auxMean[j] = runningMeanTime(auxMean[j], auxi[j], dt, elapsedTime);
        }
#endif
        if (stepcount > 1) {
            // local max of the feature variable (first occurrence in the buffer)
            if (dxb1 > 0.0 && dxb2 < 0.0) {
                realtype xThisMax = -BIG_REAL;
                int ix = 0;
                if (xb0 > xThisMax) { xThisMax = xb0; ix = 0; }
                if (xb1 > xThisMax) { xThisMax = xb1; ix = 1; }
                if (xb2 > xThisMax) { xThisMax = xb2; ix = 2; }
                timetype tThisMax = ix == 0 ? tb0 : (ix == 1 ? tb1 : tb2);
                thisNMaxima++;
                tLastMax = tThisMax;
            }
            // local min of the feature variable
            if (dxb1 < 0.0 && dxb2 > 0.0) {
                realtype xm = BIG_REAL;
                int ix = 0;
                if (xb0 < xm) { xm = xb0; ix = 0; }
                if (xb1 < xm) { xm = xb1; ix = 1; }
                if (xb2 < xm) { xm = xb2; ix = 2; }
                xLastMin = xm;
                tLastMin = ix == 0 ? tb0 : (ix == 1 ? tb1 : tb2);
            }
            if (inUpstate) {
                if (xi[CLODE_EVAR] <= xDown && dxi[CLODE_EVAR] >= dxDown) {
                    tThisDown = ti;
                    inUpstate = 0;
                    if (eventcount > 0 && eventcount <= N_STORE_EVENTS)
                        obs_list_store(io, NSCALAR, 1, eventcount - 1, tThisDown);
#if CLODE_PREC == 3
                    fVarDownstateMean = 0.0;
#else
                    fVarDownstateMean = xi[CLODE_FVAR];
#endif
                }
            } else {
#if CLODE_PREC == 3
                if (ti - tThisDown > 0.0)
                    fVarDownstateMean = fma((double)xi[CLODE_FVAR], dt, fVarDownstateMean);
#else
                realtype elapsedDownTime = (ti - tThisDown);
                if (elapsedDownTime > 0.0)
                    fVarDownstateMean = runningMeanTime(fVarDownstateMean, xi[CLODE_FVAR], dt, elapsedDownTime);
#endif
            }
        }
    }

    CLODE_DEV void finalizeFeatures(timetype ti, FeatOut &F, const ObsIO &io)
    {
        bool p = eventcount > 1;
        F.put(p ? period[0] : ZERO);
        F.put(p ? period[1] : ZERO);
        F.put(p ? period[2] : ZERO);
        F.put(p ? nMaxima[0] : ZERO);
        F.put(p ? nMaxima[1] : ZERO);
        F.put(p ? nMaxima[2] : ZERO);
        F.put(p ? upDuration[0] : ZERO);
        F.put(p ? upDuration[1] : ZERO);
        F.put(p ? upDuration[2] : ZERO);
        F.put(p ? downDuration[0] : ZERO);
        F.put(p ? downDuration[1] : ZERO);
        F.put(p ? downDuration[2] : ZERO);
        F.put(p ? duty[0] : ZERO);
        F.put(p ? duty[1] : ZERO);
        F.put(p ? duty[2] : ZERO);
        F.put(p ? activeDip[0] : ZERO);
        F.put(p ? activeDip[1] : ZERO);
        F.put(p ? activeDip[2] : ZERO);
#if CLODE_PREC == 3
        double elapsed = ti - t_start;
#endif
        for (int j = 0; j < N_VAR; ++j) {
            F.put(xMax[j]);
            F.put(xMin[j]);
#if CLODE_PREC == 3
            F.put(elapsed > 0.0 ? xMean[j] / elapsed : 0.0);
#else
            F.put(xMean[j]);
#endif
            F.put(dxMax[j]);
            F.put(dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            F.put(auxMax[j]);
            F.put(auxMin[j]);
#if CLODE_PREC == 3
            F.put(elapsed > 0.0 ? auxMean[j] / elapsed : 0.0);
#else
            F.put(auxMean[j]);
#endif
        }
        for (int j = 0; j < N_STORE_EVENTS; ++j) {
            F.put(obs_list_load(io, NSCALAR, 0, j));
            F.put(obs_list_load(io, NSCALAR, 1, j));
        }
        F.put((obsreal)eventcount);
        F.put((obsreal)stepcount);
        F.put(stepDt[0]);
        F.put(stepDt[1]);
#if CLODE_PREC == 3
        F.put(stepcount > 0 ? elapsed / (double)stepcount : 0.0);
#else
        F.put(stepDt[2]);
#endif
    }

    CLODE_DEV void finalizeObserverData(timetype ti, timetype t0)
    {
        timetype T = ti - t0;
        t_start -= T;
        tLastEvent -= T;
        tThisDown -= T;
        tLastMax -= T;
        tLastMin -= T;
        tb0 -= T;
        tb1 -= T;
        tb2 -= T;
    }
};
#endif  // thresh2

// ------------------------------------------------------------------------
#if CLODE_OBSERVER == CLODE_OBS_LOCALMAX
struct Observer {
    static constexpr int NSCALAR = clode_obs_nscalar(CLODE_OBS_LOCALMAX, N_VAR, N_AUX);
    timetype tb0, tb1, tb2;
    realtype xb0, xb1, xb2;
    realtype dxb1, dxb2;
    realtype xMax[N_VAR], xMin[N_VAR], dxMax[N_VAR], dxMin[N_VAR];
    obsreal xMean[N_VAR];
    realtype auxMax[N_AUX1], auxMin[N_AUX1];
    obsreal auxMean[N_AUX1];
    obsreal IMI[3], amp[3];
    timetype t_start, tLastMax, tLastMin;
    realtype xLastMin;
    unsigned int eventcount, stepcount;

    template <class F>
    CLODE_DEV void fields(F &f)
    {
        f.real(tb0); f.real(tb1); f.real(tb2);
        f.real(xb0); f.real(xb1); f.real(xb2);
        f.real(dxb1); f.real(dxb2);
#pragma unroll
        for (int j = 0; j < N_VAR; j++) {
            f.real(xMax[j]); f.real(xMin[j]); f.real(xMean[j]); f.real(dxMax[j]); f.real(dxMin[j]);
        }
#pragma unroll
        for (int j = 0; j < N_AUX; j++) {
            f.real(auxMax[j]); f.real(auxMin[j]); f.real(auxMean[j]);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            f.real(IMI[j]); f.real(amp[j]);
        }
        f.real(t_start); f.real(tLastMax); f.real(tLastMin); f.real(xLastMin);
        f.uint(eventcount); f.uint(stepcount);
    }

    CLODE_DEV void init(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                        const ObserverParamsR &op, const ObsIO &io)
    {
        tb2 = ti;
        xb2 = xi[CLODE_FVAR];
        dxb2 = dxi[CLODE_FVAR];
        tb0 = tb1 = ti;
        xb0 = xb1 = xb2;
        dxb1 = dxb2;
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = -BIG_REAL;
            xMin[j] = BIG_REAL;
            xMean[j] = ZERO;
            dxMax[j] = -BIG_REAL;
            dxMin[j] = BIG_REAL;
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = -BIG_REAL;
            auxMin[j] = BIG_REAL;
            auxMean[j] = ZERO;
        }
        for (int j = 0; j < N_STORE_EVENTS; ++j)
            for (int L = 0; L < 4; ++L)
                obs_list_store(io, NSCALAR, L, j, ZERO);
        IMI[0] = -BIG_REAL; IMI[1] = BIG_REAL; IMI[2] = ZERO;
        amp[0] = -BIG_REAL; amp[1] = BIG_REAL; amp[2] = ZERO;
        t_start = ti;
        tLastMax = ZERO;
        tLastMin = ZERO;
        xLastMin = BIG_REAL;
        eventcount = 0;
        stepcount = 0;
    }

    CLODE_DEV void warmup(timetype, const realtype[], const realtype[], const realtype[], const ObserverParamsR &) {}
    CLODE_DEV void initEventDetector(timetype, const realtype[], const realtype[], const realtype[],
                                     const ObserverParamsR &) {}

    CLODE_DEV bool eventFunction(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                                 const ObserverParamsR &op)
    {
        if (stepcount < 2)
            return false;
        return (dxb1 > 0.0 && dxb2 < 0.0);
    }

    CLODE_DEV bool computeEventFeatures(timetype ti, const realtype xi[], const realtype dxi[],
                                        const realtype auxi[], const ObserverParamsR &op, const ObsIO &io)
    {
        realtype xThisMax = -BIG_REAL;
        int ix = 0;
        if (xb0 > xThisMax) { xThisMax = xb0; ix = 0; }
        if (xb1 > xThisMax) { xThisMax = xb1; ix = 1; }
        if (xb2 > xThisMax) { xThisMax = xb2; ix = 2; }
        timetype tThisMax = ix == 0 ? tb0 : (ix == 1 ? tb1 : tb2);
        xThisMax = ix == 0 ? xb0 : (ix == 1 ? xb1 : xb2);
        ++eventcount;
        if (eventcount > 1) {
            obsreal thisIMI = tThisMax - tLastMax;
            IMI[0] = fmax(thisIMI, IMI[0]);
            IMI[1] = fmin(thisIMI, IMI[1]);
            runningMean(IMI[2], thisIMI, eventcount - 1);
            obsreal thisAmp = xThisMax - xLastMin;
            amp[0] = fmax(thisAmp, amp[0]);
            amp[1] = fmin(thisAmp, amp[1]);
            runningMean(amp[2], thisAmp, eventcount - 1);
        }
        tLastMax = tThisMax;
        if (eventcount <= N_STORE_EVENTS) {
            obs_list_store(io, NSCALAR, 0, eventcount - 1, tThisMax);
            obs_list_store(io, NSCALAR, 1, eventcount - 1, xThisMax);
        }
        return eventcount == op.maxEventCount;
    }

    CLODE_DEV void update(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                          const ObserverParamsR &op, const ObsIO &io)
    {
        ++stepcount;
        tb0 = tb1;
        tb1 = tb2;
        tb2 = ti;
        xb0 = xb1;
        xb1 = xb2;
        xb2 = xi[CLODE_FVAR];
        dxb1 = dxb2;
        dxb2 = dxi[CLODE_FVAR];
#if CLODE_PREC == 3
        double dt = tb2 - tb1;
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmaxf(xi[j], xMax[j]);
            xMin[j] = fminf(xi[j], xMin[j]);
            xMean[j] = fma((double)xi[j], dt, xMean[j]);
            dxMax[j] = fmaxf(dxi[j], dxMax[j]);
            dxMin[j] = fminf(dxi[j], dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = fmaxf(auxi[j], auxMax[j]);
            auxMin[j] = fminf(auxi[j], auxMin[j]);
            auxMean[j] = fma((double)auxi[j], dt, auxMean[j]);
        }
#else
        realtype dt = tb2 - tb1;
        realtype elapsedTime = ti - t_start;
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmax(xi[j], xMax[j]);
            xMin[j] = fmin(xi[j], xMin[j]);
            xMean[j] = runningMeanTime(xMean[j], xi[j], dt, elapsedTime);
            dxMax[j] = fmax(dxi[j], dxMax[j]);
            dxMin[j] = fmin(dxi[j], dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = fmax(auxi[j], auxMax[j]);
            auxMin[j] = fmin(auxi[j], auxMin[j]);
            auxMean[j] = runningMeanTime(auxMean[j], auxi[j], dt, elapsedTime);
        }
#endif
        if (stepcount < 2)
            return;
        if (dxb1 < 0.0 && dxb2 > 0.0) {
            realtype xm = BIG_REAL;
            int ix = 0;
            if (xb0 < xm) { xm = xb0; ix = 0; }
            if (xb1 < xm) { xm = xb1; ix = 1; }
            if (xb2 < xm) { xm = xb2; ix = 2; }
            tLastMin = ix == 0 ? tb0 : (ix == 1 ? tb1 : tb2);
            xLastMin = ix == 0 ? xb0 : (ix == 1 ? xb1 : xb2);
            if (eventcount > 0 && eventcount <= N_STORE_EVENTS) {
                obs_list_store(io, NSCALAR, 2, eventcount - 1, tLastMin);
                obs_list_store(io, NSCALAR, 3, eventcount - 1, xLastMin);
            }
        }
    }

    CLODE_DEV void finalizeFeatures(timetype ti, FeatOut &F, const ObsIO &io)
    {
        bool p = eventcount > 1;
        F.put(p ? IMI[0] : ZERO);
        F.put(p ? IMI[1] : ZERO);
        F.put(p ? IMI[2] : ZERO);
        F.put(p ? amp[0] : ZERO);
        F.put(p ? amp[1] : ZERO);
        F.put(p ? amp[2] : ZERO);
#if CLODE_PREC == 3
        double elapsed = ti - t_start;
#endif
        for (int j = 0; j < N_VAR; ++j) {
            F.put(xMax[j]);
            F.put(xMin[j]);
#if CLODE_PREC == 3
            F.put(elapsed > 0.0 ? xMean[j] / elapsed : 0.0);
#else
            F.put(xMean[j]);
#endif
            F.put(dxMax[j]);
            F.put(dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            F.put(auxMax[j]);
            F.put(auxMin[j]);
#if CLODE_PREC == 3
            F.put(elapsed > 0.0 ? auxMean[j] / elapsed : 0.0);
#else
            F.put(auxMean[j]);
#endif
        }
        for (int j = 0; j < N_STORE_EVENTS; ++j) {
            F.put(obs_list_load(io, NSCALAR, 0, j));
            F.put(obs_list_load(io, NSCALAR, 1, j));
            F.put(obs_list_load(io, NSCALAR, 2, j));
            F.put(obs_list_load(io, NSCALAR, 3, j));
        }
        F.put((obsreal)eventcount);
        F.put((obsreal)stepcount);
    }

    CLODE_DEV void finalizeObserverData(timetype ti, timetype t0)
    {
        timetype T = ti - t0;
        t_start -= T;
        tLastMax = tLastMax - T;
        tLastMin = tLastMin - T;
        tb0 = tb0 - T;
        tb1 = tb1 - T;
        tb2 = tb2 - T;
    }
};
#endif  // localmax

// ------------------------------------------------------------------------
#if CLODE_OBSERVER == CLODE_OBS_NHOOD2
#define CLODE_TWO_PASS 1
struct Observer {
    static constexpr int NSCALAR = clode_obs_nscalar(CLODE_OBS_NHOOD2, N_VAR, N_AUX);
    timetype tb0, tb1, tb2;
    realtype xbe1, xbe2;  // event-variable x buffer
    realtype dxb1, dxb2;  // feature-variable dx buffer
    realtype x0[N_VAR], xMax[N_VAR], xMin[N_VAR], xRange[N_VAR], dxMax[N_VAR], dxMin[N_VAR];
    obsreal xMean[N_VAR];
    realtype auxMax[N_AUX1], auxMin[N_AUX1];
    obsreal auxMean[N_AUX1];
    obsreal nMaxima[3], period[3], stepDt[3];
    timetype t_start, tLastEvent;
    realtype xThreshold;
    unsigned int thisNMaxima, foundX0, isInNhood, eventcount, stepcount;

    template <class F>
    CLODE_DEV void fields(F &f)
    {
        f.real(tb0); f.real(tb1); f.real(tb2);
        f.real(xbe1); f.real(xbe2);
        f.real(dxb1); f.real(dxb2);
#pragma unroll
        for (int j = 0; j < N_VAR; j++) {
            f.real(x0[j]); f.real(xMax[j]); f.real(xMin[j]); f.real(xMean[j]);
            f.real(xRange[j]); f.real(dxMax[j]); f.real(dxMin[j]);
        }
#pragma unroll
        for (int j = 0; j < N_AUX; j++) {
            f.real(auxMax[j]); f.real(auxMin[j]); f.real(auxMean[j]);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            f.real(nMaxima[j]); f.real(period[j]); f.real(stepDt[j]);
        }
        f.real(t_start); f.real(tLastEvent); f.real(xThreshold);
        f.uint(thisNMaxima); f.uint(foundX0); f.uint(isInNhood); f.uint(eventcount); f.uint(stepcount);
    }

    CLODE_DEV void init(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                        const ObserverParamsR &op, const ObsIO &io)
    {
        tb2 = ti;
        tb0 = tb1 = ti;
        xbe2 = xi[CLODE_EVAR];
        xbe1 = xbe2;
        dxb2 = dxi[CLODE_FVAR];
        dxb1 = dxb2;
        for (int j = 0; j < N_VAR; ++j) {
            x0[j] = xi[j];
            xMax[j] = -BIG_REAL;
            xMin[j] = BIG_REAL;
            xMean[j] = ZERO;
            xRange[j] = ZERO;
            dxMax[j] = -BIG_REAL;
            dxMin[j] = BIG_REAL;
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = -BIG_REAL;
            auxMin[j] = BIG_REAL;
            auxMean[j] = ZERO;
        }
        for (int j = 0; j < N_STORE_EVENTS; ++j)
            obs_list_store(io, NSCALAR, 0, j, ZERO);
        nMaxima[0] = -BIG_REAL; nMaxima[1] = BIG_REAL; nMaxima[2] = ZERO;
        period[0] = -BIG_REAL; period[1] = BIG_REAL; period[2] = ZERO;
        stepDt[0] = -BIG_REAL; stepDt[1] = BIG_REAL; stepDt[2] = ZERO;
        t_start = ti;
        tLastEvent = ZERO;
        xThreshold = ZERO;
        thisNMaxima = 0;
        foundX0 = 0;
        isInNhood = 0;
        eventcount = 0;
        stepcount = 0;
    }

    CLODE_DEV void warmup(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                          const ObserverParamsR &op)
    {
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmax(xi[j], xMax[j]);
            xMin[j] = fmin(xi[j], xMin[j]);
        }
    }

    CLODE_DEV void initEventDetector(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                                     const ObserverParamsR &op)
    {
        for (int j = 0; j < N_VAR; ++j)
            xRange[j] = xMax[j] - xMin[j];
        xThreshold = xMin[CLODE_EVAR] + op.xDownThresh * xRange[CLODE_EVAR];
    }

    CLODE_DEV bool eventFunction(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                                 const ObserverParamsR &op)
    {
        if (stepcount < 2 || !foundX0)
            return false;
        realtype thisXdiff[N_VAR];
        for (int j = 0; j < N_VAR; ++j)
            thisXdiff[j] = (xi[j] - x0[j]) / xRange[j];
        unsigned int lastInNhood = isInNhood;
        realtype result = ZERO;
        for (int k = 0; k < N_VAR; k++)
            result += thisXdiff[k] * thisXdiff[k];
        isInNhood = sqrt(result) < op.nHoodRadius;
        if (lastInNhood && !isInNhood)
            return true;
        return false;
    }

    CLODE_DEV bool computeEventFeatures(timetype ti, const realtype xi[], const realtype dxi[],
                                        const realtype auxi[], const ObserverParamsR &op, const ObsIO &io)
    {
        ++eventcount;
        timetype tThisEvent = ti;
        if (eventcount > 1) {
            nMaxima[0] = fmax((obsreal)thisNMaxima, nMaxima[0]);
            nMaxima[1] = fmin((obsreal)thisNMaxima, nMaxima[1]);
            runningMean(nMaxima[2], (obsreal)thisNMaxima, eventcount - 1);
            obsreal thisPeriod = tThisEvent - tLastEvent;
            period[0] = fmax(thisPeriod, period[0]);
            period[1] = fmin(thisPeriod, period[1]);
            runningMean(period[2], thisPeriod, eventcount - 1);
        }
        tLastEvent = tThisEvent;
        thisNMaxima = 0;
        if (eventcount <= N_STORE_EVENTS)
            obs_list_store(io, NSCALAR, 0, eventcount - 1, tThisEvent);
        return eventcount == op.maxEventCount;
    }

    CLODE_DEV void update(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                          const ObserverParamsR &op, const ObsIO &io)
    {
        ++stepcount;
        tb0 = tb1;
        tb1 = tb2;
        tb2 = ti;
        xbe1 = xbe2;
        xbe2 = xi[CLODE_EVAR];
        dxb1 = dxb2;
        dxb2 = dxi[CLODE_FVAR];
#if CLODE_PREC == 3
        double dt = tb2 - tb1;
        stepDt[0] = fmax(dt, stepDt[0]);
        stepDt[1] = fmin(dt, stepDt[1]);
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmaxf(xi[j], xMax[j]);
            xMin[j] = fminf(xi[j], xMin[j]);
            xMean[j] = fma((double)xi[j], dt, xMean[j]);
            dxMax[j] = fmaxf(dxi[j], dxMax[j]);
            dxMin[j] = fminf(dxi[j], dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = fmaxf(auxi[j], auxMax[j]);
            auxMin[j] = fminf(auxi[j], auxMin[j]);
            auxMean[j] = fma((double)auxi[j], dt, auxMean[j]);
        }
#else
        realtype dt = tb2 - tb1;
        realtype elapsedTime = ti - t_start;
        stepDt[0] = fmax(dt, stepDt[0]);
        stepDt[1] = fmin(dt, stepDt[1]);
        runningMean(stepDt[2], dt, stepcount);
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmax(xi[j], xMax[j]);
            xMin[j] = fmin(xi[j], xMin[j]);
            xMean[j] = runningMeanTime(xMean[j], xi[j], dt, elapsedTime);
            dxMax[j] = fmax(dxi[j], dxMax[j]);
            dxMin[j] = fmin(dxi[j], dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = fmax(auxi[j], auxMax[j]);
            auxMin[j] = fmin(auxi[j], auxMin[j]);
            auxMean[j] = runningMeanTime(auxMean[j], auxi[j], dt, elapsedTime);
        }
#endif
        if (stepcount < 2)
            return;
        if (foundX0) {
            if (dxb1 >= 0.0 && dxb2 < 0.0)
                thisNMaxima++;
            return;
        }
        if (xbe1 > xThreshold && xbe2 < xThreshold) {
            foundX0 = 1;
            isInNhood = 1;
            for (int j = 0; j < N_VAR; ++j)
                x0[j] = xi[j];
        }
    }

    CLODE_DEV void finalizeFeatures(timetype ti, FeatOut &F, const ObsIO &io)
    {
        bool p = eventcount > 1;
        F.put(p ? period[0] : ZERO);
        F.put(p ? period[1] : ZERO);
        F.put(p ? period[2] : ZERO);
        F.put(p ? nMaxima[0] : ZERO);
        F.put(p ? nMaxima[1] : ZERO);
        F.put(p ? nMaxima[2] : ZERO);
#if CLODE_PREC == 3
        double elapsed = ti - t_start;
#endif
        for (int j = 0; j < N_VAR; ++j) {
            F.put(xMax[j]);
            F.put(xMin[j]);
#if CLODE_PREC == 3
            F.put(elapsed > 0.0 ? xMean[j] / elapsed : 0.0);
#else
            F.put(xMean[j]);
#endif
            F.put(xRange[j]);
            F.put(x0[j]);
            F.put(dxMax[j]);
            F.put(dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            F.put(auxMax[j]);
            F.put(auxMin[j]);
#if CLODE_PREC == 3
            F.put(elapsed > 0.0 ? auxMean[j] / elapsed : 0.0);
#else
            F.put(auxMean[j]);
#endif
        }
        for (int j = 0; j < N_STORE_EVENTS; ++j)
            F.put(obs_list_load(io, NSCALAR, 0, j));
        F.put((obsreal)eventcount);
        F.put((obsreal)stepcount);
        F.put(stepDt[0]);
        F.put(stepDt[1]);
#if CLODE_PREC == 3
        F.put(stepcount > 0 ? elapsed / (double)stepcount : 0.0);
#else
        F.put(stepDt[2]);
#endif
    }

    CLODE_DEV void finalizeObserverData(timetype ti, timetype t0)
    {
        timetype T = ti - t0;
        t_start -= T;
        tLastEvent -= T;
        tb0 -= T;
        tb1 -= T;
        tb2 -= T;
    }
};
#endif  // nhood2

// ------------------------------------------------------------------------
// basic (observer_basic.clh:17-90) and basicall (observer_basic_allVar.clh:30-124):
// one-pass trajectory extents and time means, no events.  basic tracks the
// feature variable only.  The time step is measured from the previous update
// (t_last), which the continuation shift leaves unchanged, as in the reference.
#if CLODE_OBSERVER == CLODE_OBS_BASIC || CLODE_OBSERVER == CLODE_OBS_BASICALL
#if CLODE_OBSERVER == CLODE_OBS_BASIC
#define CLODE_BASIC_NV 1
#define CLODE_BASIC_NA 0
#define CLODE_BASIC_VAR(j) CLODE_FVAR
#else
#define CLODE_BASIC_NV N_VAR
#define CLODE_BASIC_NA N_AUX
#define CLODE_BASIC_VAR(j) (j)
#endif
struct Observer {
    static constexpr int NSCALAR = clode_obs_nscalar(CLODE_OBSERVER, N_VAR, N_AUX);
    realtype xMax[CLODE_BASIC_NV], xMin[CLODE_BASIC_NV], dxMax[CLODE_BASIC_NV], dxMin[CLODE_BASIC_NV];
    obsreal xMean[CLODE_BASIC_NV];  // running time mean (plain) or integral (fp32c)
    realtype auxMax[N_AUX1], auxMin[N_AUX1];
    obsreal auxMean[N_AUX1];
    timetype t_last, t_start;
    unsigned int stepcount;

    template <class F>
    CLODE_DEV void fields(F &f)
    {
#pragma unroll
        for (int j = 0; j < CLODE_BASIC_NV; j++) {
            f.real(xMax[j]); f.real(xMin[j]); f.real(xMean[j]); f.real(dxMax[j]); f.real(dxMin[j]);
        }
#pragma unroll
        for (int j = 0; j < CLODE_BASIC_NA; j++) {
            f.real(auxMax[j]); f.real(auxMin[j]); f.real(auxMean[j]);
        }
        f.real(t_last); f.real(t_start);
        f.uint(stepcount);
    }

    CLODE_DEV void init(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                        const ObserverParamsR &op, const ObsIO &io)
    {
        for (int j = 0; j < CLODE_BASIC_NV; ++j) {
            xMax[j] = -BIG_REAL;
            xMin[j] = BIG_REAL;
            xMean[j] = ZERO;
            dxMax[j] = -BIG_REAL;
            dxMin[j] = BIG_REAL;
        }
        for (int j = 0; j < CLODE_BASIC_NA; ++j) {
            auxMax[j] = -BIG_REAL;
            auxMin[j] = BIG_REAL;
            auxMean[j] = ZERO;
        }
        t_last = ti;
        t_start = ti;
        stepcount = 0;
    }

    CLODE_DEV void warmup(timetype, const realtype[], const realtype[], const realtype[], const ObserverParamsR &) {}
    CLODE_DEV void initEventDetector(timetype, const realtype[], const realtype[], const realtype[],
                                     const ObserverParamsR &) {}
    CLODE_DEV bool eventFunction(timetype, const realtype[], const realtype[], const realtype[],
                                 const ObserverParamsR &)
    {
        return false;
    }
    CLODE_DEV bool computeEventFeatures(timetype, const realtype[], const realtype[], const realtype[],
                                        const ObserverParamsR &, const ObsIO &)
    {
        return false;
    }

    CLODE_DEV void update(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                          const ObserverParamsR &op, const ObsIO &io)
    {
        ++stepcount;
#if CLODE_PREC == 3
        double dt = ti - t_last;
        t_last = ti;
        for (int j = 0; j < CLODE_BASIC_NV; ++j) {
            const int v = CLODE_BASIC_VAR(j);
            xMax[j] = fmaxf(xi[v], xMax[j]);
            xMin[j] = fminf(xi[v], xMin[j]);
            xMean[j] = fma((double)xi[v], dt, xMean[j]);
            dxMax[j] = fmaxf(dxi[v], dxMax[j]);
            dxMin[j] = fminf(dxi[v], dxMin[j]);
        }
        for (int j = 0; j < CLODE_BASIC_NA; ++j) {
            auxMax[j] = fmaxf(auxi[j], auxMax[j]);
            auxMin[j] = fminf(auxi[j], auxMin[j]);
            auxMean[j] = fma((double)auxi[j], dt, auxMean[j]);
        }
#else
        realtype dt = ti - t_last;
        t_last = ti;
        realtype elapsedTime = ti - t_start;
        for (int j = 0; j < CLODE_BASIC_NV; ++j) {
            const int v = CLODE_BASIC_VAR(j);
            xMax[j] = fmax(xi[v], xMax[j]);
            xMin[j] = fmin(xi[v], xMin[j]);
            xMean[j] = runningMeanTime(xMean[j], xi[v], dt, elapsedTime);
            dxMax[j] = fmax(dxi[v], dxMax[j]);
            dxMin[j] = fmin(dxi[v], dxMin[j]);
        }
        for (int j = 0; j < CLODE_BASIC_NA; ++j) {
            auxMax[j] = fmax(auxi[j], auxMax[j]);
            auxMin[j] = fmin(auxi[j], auxMin[j]);
            auxMean[j] = runningMeanTime(auxMean[j], auxi[j], dt, elapsedTime);
        }
#endif
    }

    CLODE_DEV void finalizeFeatures(timetype ti, FeatOut &F, const ObsIO &io)
    {
#if CLODE_PREC == 3
        double elapsed = ti - t_start;
#endif
        for (int j = 0; j < CLODE_BASIC_NV; ++j) {
            F.put(xMax[j]);
            F.put(xMin[j]);
#if CLODE_PREC == 3
            F.put(elapsed != 0.0 ? xMean[j] / elapsed : 0.0);
#else
            F.put(xMean[j]);
#endif
            F.put(dxMax[j]);
            F.put(dxMin[j]);
        }
        for (int j = 0; j < CLODE_BASIC_NA; ++j) {
            F.put(auxMax[j]);
            F.put(auxMin[j]);
#if CLODE_PREC == 3
            F.put(elapsed != 0.0 ? auxMean[j] / elapsed : 0.0);
#else
            F.put(auxMean[j]);
#endif
        }
        F.put((obsreal)stepcount);
    }

    CLODE_DEV void finalizeObserverData(timetype ti, timetype t0)
    {
        timetype T = ti - t0;
#if CLODE_PREC == 3
        // The next update measures its step from the unshifted t_last = ti,
        // so the running mean's recurrence m' E' = m (E' - dt') + x dt' weighs
        // the mean so far by E' - dt' = E + T (E = ti - t_start, the time
        // covered so far), not by E.  The integral (m E) takes the same
        // factor, so that it keeps tracking m' E'.
        const double E = ti - t_start;
        if (E > 0.0) {
            const double s = (E + T) / E;
            for (int j = 0; j < CLODE_BASIC_NV; ++j)
                xMean[j] *= s;
            for (int j = 0; j < CLODE_BASIC_NA; ++j)
                auxMean[j] *= s;
        }
#endif
        t_start -= T;
    }
};
#endif  // basic, basicall

// ------------------------------------------------------------------------
// nhood1 (observer_neighborhood_1.clh:40-310): one-pass neighbourhood return.
// The centre is the state at the first local minimum of the event variable;
// an event is the entry into the L2 ball of radius nHoodRadius in
// range-normalised coordinates (ranges from the running extents).  Means are
// step-count means.  Of the 3-point buffers, x is kept for every variable (the
// centre takes all components), dx only for the event and feature variables.
#if CLODE_OBSERVER == CLODE_OBS_NHOOD1
struct Observer {
    static constexpr int NSCALAR = clode_obs_nscalar(CLODE_OBS_NHOOD1, N_VAR, N_AUX);
    timetype tb0, tb1, tb2;
    realtype xb0[N_VAR], xb1[N_VAR], xb2[N_VAR];
    realtype dxe1, dxe2, dxf1, dxf2;
    realtype x0[N_VAR], xMax[N_VAR], xMin[N_VAR], dxMax[N_VAR], dxMin[N_VAR];
    obsreal xMean[N_VAR];  // running mean (plain) or sum (fp32c)
    realtype auxMax[N_AUX1], auxMin[N_AUX1];
    obsreal auxMean[N_AUX1];
    obsreal nMaxima[3], period[3], stepDt[3];
    timetype t_start, tLastEvent;
    unsigned int thisNMaxima, eventcount, stepcount, foundX0, isInNhood;

    template <class F>
    CLODE_DEV void fields(F &f)
    {
        f.real(tb0); f.real(tb1); f.real(tb2);
#pragma unroll
        for (int j = 0; j < N_VAR; j++) {
            f.real(xb0[j]); f.real(xb1[j]); f.real(xb2[j]);
        }
        f.real(dxe1); f.real(dxe2); f.real(dxf1); f.real(dxf2);
#pragma unroll
        for (int j = 0; j < N_VAR; j++) {
            f.real(x0[j]); f.real(xMax[j]); f.real(xMin[j]); f.real(xMean[j]); f.real(dxMax[j]); f.real(dxMin[j]);
        }
#pragma unroll
        for (int j = 0; j < N_AUX; j++) {
            f.real(auxMax[j]); f.real(auxMin[j]); f.real(auxMean[j]);
        }
#pragma unroll
        for (int j = 0; j < 3; j++) {
            f.real(nMaxima[j]); f.real(period[j]); f.real(stepDt[j]);
        }
        f.real(t_start); f.real(tLastEvent);
        f.uint(thisNMaxima); f.uint(eventcount); f.uint(stepcount); f.uint(foundX0); f.uint(isInNhood);
    }

    CLODE_DEV void init(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                        const ObserverParamsR &op, const ObsIO &io)
    {
        tb0 = tb1 = tb2 = ti;
        for (int j = 0; j < N_VAR; ++j) {
            xb0[j] = xb1[j] = xb2[j] = xi[j];
            x0[j] = xi[j];
            xMax[j] = -BIG_REAL;
            xMin[j] = BIG_REAL;
            xMean[j] = ZERO;
            dxMax[j] = -BIG_REAL;
            dxMin[j] = BIG_REAL;
        }
        dxe1 = dxe2 = dxi[CLODE_EVAR];
        dxf1 = dxf2 = dxi[CLODE_FVAR];
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = -BIG_REAL;
            auxMin[j] = BIG_REAL;
            auxMean[j] = ZERO;
        }
        nMaxima[0] = -BIG_REAL; nMaxima[1] = BIG_REAL; nMaxima[2] = ZERO;
        period[0] = -BIG_REAL; period[1] = BIG_REAL; period[2] = ZERO;
        stepDt[0] = -BIG_REAL; stepDt[1] = BIG_REAL; stepDt[2] = ZERO;
        t_start = ti;
        tLastEvent = ZERO;
        thisNMaxima = 0;
        eventcount = 0;
        stepcount = 0;
        foundX0 = 0;
        isInNhood = 0;
    }

    CLODE_DEV void warmup(timetype, const realtype[], const realtype[], const realtype[], const ObserverParamsR &) {}
    CLODE_DEV void initEventDetector(timetype, const realtype[], const realtype[], const realtype[],
                                     const ObserverParamsR &) {}

    CLODE_DEV bool eventFunction(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                                 const ObserverParamsR &op)
    {
        if (stepcount < 2 || !foundX0)
            return false;
        if (xMax[CLODE_FVAR] - xMin[CLODE_FVAR] < op.minXamp)
            return false;
        unsigned int lastInNhood = isInNhood;
        realtype thisXdiff[N_VAR];
        for (int j = 0; j < N_VAR; ++j)
            thisXdiff[j] = fabs(xi[j] - x0[j]) / (xMax[j] - xMin[j]);
        realtype result = ZERO;
        for (int k = 0; k < N_VAR; k++)
            result += thisXdiff[k] * thisXdiff[k];
        isInNhood = sqrt(result) <= op.nHoodRadius;
        return isInNhood && !lastInNhood;
    }

    CLODE_DEV bool computeEventFeatures(timetype ti, const realtype xi[], const realtype dxi[],
                                        const realtype auxi[], const ObserverParamsR &op, const ObsIO &io)
    {
        ++eventcount;
        timetype tThisEvent = ti;
        if (eventcount > 1) {
            nMaxima[0] = fmax((obsreal)thisNMaxima, nMaxima[0]);
            nMaxima[1] = fmin((obsreal)thisNMaxima, nMaxima[1]);
            runningMean(nMaxima[2], (obsreal)thisNMaxima, eventcount - 1);
            obsreal thisPeriod = tThisEvent - tLastEvent;
            period[0] = fmax(thisPeriod, period[0]);
            period[1] = fmin(thisPeriod, period[1]);
            runningMean(period[2], thisPeriod, eventcount - 1);
        }
        tLastEvent = tThisEvent;
        thisNMaxima = 0;
        return eventcount >= op.maxEventCount;
    }

    CLODE_DEV void update(timetype ti, const realtype xi[], const realtype dxi[], const realtype auxi[],
                          const ObserverParamsR &op, const ObsIO &io)
    {
        ++stepcount;
        tb0 = tb1;
        tb1 = tb2;
        tb2 = ti;
        for (int j = 0; j < N_VAR; ++j) {
            xb0[j] = xb1[j];
            xb1[j] = xb2[j];
            xb2[j] = xi[j];
        }
        dxe1 = dxe2;
        dxe2 = dxi[CLODE_EVAR];
        dxf1 = dxf2;
        dxf2 = dxi[CLODE_FVAR];
#if CLODE_PREC == 3
        double thisDt = tb2 - tb1;
        stepDt[0] = fmax(thisDt, stepDt[0]);
        stepDt[1] = fmin(thisDt, stepDt[1]);
        stepDt[2] += thisDt;
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmaxf(xi[j], xMax[j]);
            xMin[j] = fminf(xi[j], xMin[j]);
            xMean[j] += (double)xi[j];
            dxMax[j] = fmaxf(dxi[j], dxMax[j]);
            dxMin[j] = fminf(dxi[j], dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = fmaxf(auxi[j], auxMax[j]);
            auxMin[j] = fminf(auxi[j], auxMin[j]);
            auxMean[j] += (double)auxi[j];
        }
#else
        realtype thisDt = tb2 - tb1;
        stepDt[0] = fmax(thisDt, stepDt[0]);
        stepDt[1] = fmin(thisDt, stepDt[1]);
        runningMean(stepDt[2], thisDt, stepcount);
        for (int j = 0; j < N_VAR; ++j) {
            xMax[j] = fmax(xi[j], xMax[j]);
            xMin[j] = fmin(xi[j], xMin[j]);
            runningMean(xMean[j], xi[j], stepcount);
            dxMax[j] = fmax(dxi[j], dxMax[j]);
            dxMin[j] = fmin(dxi[j], dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            auxMax[j] = fmax(auxi[j], auxMax[j]);
            auxMin[j] = fmin(auxi[j], auxMin[j]);
            runningMean(auxMean[j], auxi[j], stepcount);
        }
#endif
        if (stepcount < 2)
            return;
        if (!foundX0) {
            if (dxe1 <= 0.0 && dxe2 > 0.0) {
                // first-occurrence minimum of the event variable's buffer
                realtype xm = BIG_REAL;
                int ix = 0;
                if (xb0[CLODE_EVAR] < xm) { xm = xb0[CLODE_EVAR]; ix = 0; }
                if (xb1[CLODE_EVAR] < xm) { xm = xb1[CLODE_EVAR]; ix = 1; }
                if (xb2[CLODE_EVAR] < xm) { xm = xb2[CLODE_EVAR]; ix = 2; }
                tLastEvent = ix == 0 ? tb0 : (ix == 1 ? tb1 : tb2);
                foundX0 = 1;
                for (int j = 0; j < N_VAR; ++j)
                    x0[j] = ix == 0 ? xb0[j] : (ix == 1 ? xb1[j] : xb2[j]);
            }
        } else if (dxf1 >= 0.0 && dxf2 < 0.0) {
            thisNMaxima++;
        }
    }

    CLODE_DEV void finalizeFeatures(timetype ti, FeatOut &F, const ObsIO &io)
    {
        bool p = eventcount > 1;
        F.put(p ? period[0] : ZERO);
        F.put(p ? period[1] : ZERO);
        F.put(p ? period[2] : ZERO);
        F.put(p ? nMaxima[0] : ZERO);
        F.put(p ? nMaxima[1] : ZERO);
        F.put(p ? nMaxima[2] : ZERO);
#if CLODE_PREC == 3
        const double inv = stepcount > 0 ? 1.0 / (double)stepcount : 0.0;
#endif
        for (int j = 0; j < N_VAR; ++j) {
            F.put(xMax[j]);
            F.put(xMin[j]);
#if CLODE_PREC == 3
            F.put(xMean[j] * inv);
#else
            F.put(xMean[j]);
#endif
            F.put(dxMax[j]);
            F.put(dxMin[j]);
        }
        for (int j = 0; j < N_AUX; ++j) {
            F.put(auxMax[j]);
            F.put(auxMin[j]);
#if CLODE_PREC == 3
            F.put(auxMean[j] * inv);
#else
            F.put(auxMean[j]);
#endif
        }
        F.put((obsreal)(unsigned int)(eventcount - 1u));  // periods; wraps at eventcount 0 as in the reference
        F.put((obsreal)stepcount);
        F.put(stepDt[0]);
        F.put(stepDt[1]);
#if CLODE_PREC == 3
        F.put(stepDt[2] * inv);
#else
        F.put(stepDt[2]);
#endif
    }

    CLODE_DEV void finalizeObserverData(timetype ti, timetype t0)
    {
        timetype T = ti - t0;
        t_start -= T;
        tLastEvent -= T;
        tb0 -= T;
        tb1 -= T;
        tb2 -= T;
    }
};
#endif  // nhood1

#endif  // CLODE_OBSERVERS_H
