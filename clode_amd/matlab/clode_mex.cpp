// MATLAB gateways over the C ABI (include/clode_hip.h).
//
// One source, three gateways, selected at compile time with CLODE_MEX_KIND:
//   0  clODEmex            base simulator      (matlab/clODEmex.cpp)
//   1  clODEfeaturesmex    feature simulator   (matlab/clODEfeaturesmex.cpp)
//   2  clODEtrajectorymex  trajectory simulator (matlab/clODEtrajectorymex.cpp)
// They speak the reference's action-string protocol, so the reference's
// MATLAB classes (matlab/clODE.m, clODEfeatures.m, clODEtrajectory.m) drive
// them unchanged:  h = clODEmex('new', prob, stepper, single, platformID,
// deviceID[, observer]);  clODEmex('settspan', h, [t0 t1]);  ...
//   * instances live in a table keyed by a positive integer handle, and the
//     MEX file is locked while any instance exists (matlab/clODEmex.cpp:91-94);
//   * the problem struct has fields clRHSfilename, nVar, nPar, nAux, nWiener,
//     varNames, parNames, auxNames (matlab/clODEmexHelpers.hpp:46-70); the
//     solver struct dt, dtmax, abstol, reltol, max_steps, max_store, nout
//     (:33-43); the observer struct eVarIx, fVarIx (1-based), maxEventCount,
//     minXamp, minIMI, nHoodRadius, xUpThresh, xDownThresh, dxUpThresh,
//     dxDownThresh, eps_dx and, if present, maxEventTimestamps
//     (matlab/clODEfeaturesmex.cpp:431-444);
//   * arrays cross as double column vectors in the C ABI's layouts.
// The reference gateways still call two members its core no longer has,
// initialize(tspan, x0, pars, sp) and setNpts(n) (clODEmex.cpp:215-230);
// here 'initialize' sets the span, problem data and solver parameters, and
// 'setnpts' sets zero-filled problem data of n members.
#include <cctype>
#include <map>
#include <string>
#include <vector>

#include "clode_hip.h"
#include "mex.h"

#ifndef CLODE_MEX_KIND
#define CLODE_MEX_KIND 0
#endif

namespace {

typedef std::map<unsigned int, clode_sim *> Table;

Table &table()
{
    static Table t;
    return t;
}

void fail(const std::string &msg) { mexErrMsgTxt(msg.c_str()); }

void check(int rc)
{
    if (rc != CLODE_OK)
        fail(std::string("clode: ") + clode_last_error());
}

std::string str(const mxArray *a)
{
    if (!a || !mxIsChar(a))
        fail("expected a character array");
    char *c = mxArrayToString(a);
    std::string s(c ? c : "");
    mxFree(c);
    return s;
}

double scalar(const mxArray *a, const char *what)
{
    if (!a || mxGetNumberOfElements(a) < 1)
        fail(std::string("missing ") + what);
    return mxGetScalar(a);
}

const mxArray *field(const mxArray *s, const char *name, bool required = true)
{
    const mxArray *f = mxIsStruct(s) ? mxGetField(s, 0, name) : nullptr;
    if (!f && required)
        fail(std::string("struct field '") + name + "' is missing");
    return f;
}

std::vector<double> vec(const mxArray *a)
{
    if (!a || !mxIsDouble(a))
        fail("expected a double array");
    const double *p = mxGetPr(a);
    return std::vector<double>(p, p + mxGetNumberOfElements(a));
}

std::string names(const mxArray *cell)
{
    std::string out;
    if (!cell)
        return out;
    if (!mxIsCell(cell))
        fail("names must be a cell array of strings");
    for (size_t i = 0; i < mxGetNumberOfElements(cell); ++i) {
        if (i)
            out += "\n";
        out += str(mxGetCell(cell, i));
    }
    return out;
}

// problem struct -> clode_problem_info (strings kept alive by the caller)
struct Problem {
    std::string file, vars, pars, aux;
    clode_problem_info pi;
};

void problem(const mxArray *s, Problem &p)
{
    if (!mxIsStruct(s))
        fail("the problem must be a struct");
    p.file = str(field(s, "clRHSfilename"));
    p.vars = names(field(s, "varNames", false));
    p.pars = names(field(s, "parNames", false));
    p.aux = names(field(s, "auxNames", false));
    p.pi.src_file = p.file.c_str();
    p.pi.num_var = (int)scalar(field(s, "nVar"), "nVar");
    p.pi.num_par = (int)scalar(field(s, "nPar"), "nPar");
    p.pi.num_aux = (int)scalar(field(s, "nAux"), "nAux");
    p.pi.num_noise = (int)scalar(field(s, "nWiener"), "nWiener");
    p.pi.var_names = p.vars.c_str();
    p.pi.par_names = p.pars.c_str();
    p.pi.aux_names = p.aux.c_str();
}

clode_solver_params solver(const mxArray *s)
{
    clode_solver_params sp;
    sp.dt = scalar(field(s, "dt"), "dt");
    sp.dtmax = scalar(field(s, "dtmax"), "dtmax");
    sp.abstol = scalar(field(s, "abstol"), "abstol");
    sp.reltol = scalar(field(s, "reltol"), "reltol");
    sp.max_steps = (unsigned int)scalar(field(s, "max_steps"), "max_steps");
    sp.max_store = (unsigned int)scalar(field(s, "max_store"), "max_store");
    sp.nout = (unsigned int)scalar(field(s, "nout"), "nout");
    return sp;
}

[[maybe_unused]] clode_observer_params observer(const mxArray *s)
{
    clode_observer_params op;
    const double e = scalar(field(s, "eVarIx"), "eVarIx"), f = scalar(field(s, "fVarIx"), "fVarIx");
    if (e < 1 || f < 1)
        fail("eVarIx and fVarIx are 1-based");
    op.eVarIx = (unsigned int)e - 1;
    op.fVarIx = (unsigned int)f - 1;
    op.maxEventCount = (unsigned int)scalar(field(s, "maxEventCount"), "maxEventCount");
    const mxArray *ts = field(s, "maxEventTimestamps", false);
    op.maxEventTimestamps = ts ? (unsigned int)scalar(ts, "maxEventTimestamps") : 0u;
    op.minXamp = scalar(field(s, "minXamp"), "minXamp");
    op.minIMI = scalar(field(s, "minIMI"), "minIMI");
    op.nHoodRadius = scalar(field(s, "nHoodRadius"), "nHoodRadius");
    op.xUpThresh = scalar(field(s, "xUpThresh"), "xUpThresh");
    op.xDownThresh = scalar(field(s, "xDownThresh"), "xDownThresh");
    op.dxUpThresh = scalar(field(s, "dxUpThresh"), "dxUpThresh");
    op.dxDownThresh = scalar(field(s, "dxDownThresh"), "dxDownThresh");
    op.eps_dx = scalar(field(s, "eps_dx"), "eps_dx");
    return op;
}

[[maybe_unused]] mxArray *column(const std::vector<double> &v)
{
    mxArray *a = mxCreateDoubleMatrix(v.size(), 1, mxREAL);
    double *p = mxGetPr(a);
    for (size_t i = 0; i < v.size(); ++i)
        p[i] = v[i];
    return a;
}

mxArray *row(const std::vector<double> &v)
{
    mxArray *a = mxCreateDoubleMatrix(1, v.size(), mxREAL);
    double *p = mxGetPr(a);
    for (size_t i = 0; i < v.size(); ++i)
        p[i] = v[i];
    return a;
}

// full string from a C ABI string getter (*needed = length + 1)
template <class Getter>
std::string get_string(Getter get)
{
    size_t need = 0;
    check(get(nullptr, 0, &need));
    std::string buf(need > 0 ? need : 1, '\0');
    check(get(&buf[0], buf.size(), &need));
    buf.resize(need > 0 ? need - 1 : 0);
    return buf;
}

// '\n'-separated name list from a C ABI string getter -> cell column
template <class Getter>
mxArray *name_cell(Getter get)
{
    const std::string buf = get_string(get);
    std::vector<std::string> parts;
    std::string cur;
    for (char c : buf) {
        if (c == '\n') {
            parts.push_back(cur);
            cur.clear();
        } else {
            cur += c;
        }
    }
    if (!buf.empty())
        parts.push_back(cur);
    mxArray *cell = mxCreateCellMatrix(parts.size(), 1);
    for (size_t i = 0; i < parts.size(); ++i)
        mxSetCell(cell, i, mxCreateString(parts[i].c_str()));
    return cell;
}

size_t members(clode_sim *s)
{
    int n = 0;
    check(clode_get_n_pts(s, &n));
    return (size_t)n;
}

clode_sim *instance(int nrhs, const mxArray *prhs[])
{
    if (nrhs < 2 || mxGetNumberOfElements(prhs[1]) != 1)
        fail("Specify an instance with an integer handle.");
    const unsigned int h = (unsigned int)mxGetScalar(prhs[1]);
    auto it = table().find(h);
    if (it == table().end())
        fail("No instance corresponding to handle " + std::to_string(h) + " found.");
    return it->second;
}

void need(int nrhs, int n, const std::string &action)
{
    if (nrhs < n)
        fail("too few arguments for '" + action + "'");
}

}  // namespace

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs < 1 || !mxIsChar(prhs[0]))
        fail("First input must be an action string ('new', 'delete', or a method name).");
    std::string action = str(prhs[0]);
    for (auto &c : action)
        c = (char)std::tolower((unsigned char)c);

    if (action == "new") {
        // ('new', prob, stepper, single, platformID, deviceID[, observer])
        need(nrhs, CLODE_MEX_KIND == 1 ? 7 : 6, action);
        Problem p;
        problem(prhs[1], p);
        const std::string stepper = str(prhs[2]);
        const int precision = scalar(prhs[3], "precision") != 0 ? CLODE_PREC_SINGLE : CLODE_PREC_DOUBLE;
        const int device = (int)scalar(prhs[5], "deviceID");
        std::string obs = CLODE_MEX_KIND == 1 ? str(prhs[6]) : std::string();
        clode_observer_params op;
        const clode_observer_params *opp = nullptr;
        if (CLODE_MEX_KIND == 1) {
            // the reference's default observer parameters (CLODEpython.cpp:300-311)
            op = clode_observer_params{0, 0, 100, 0, 0.0, 0.0, 0.05, 0.2, 0.2, 0.0, 0.0, 0.0};
            opp = &op;
        }
        clode_sim *s = nullptr;
        check(clode_create(&s, CLODE_MEX_KIND, &p.pi, stepper.c_str(), CLODE_MEX_KIND == 1 ? obs.c_str() : nullptr,
                           opp, precision, device, nullptr));
        const unsigned int h = table().empty() ? 1u : table().rbegin()->first + 1u;
        table()[h] = s;
        mexLock();
        plhs[0] = mxCreateDoubleScalar((double)h);
        return;
    }
    if (action == "delete") {
        clode_sim *s = instance(nrhs, prhs);
        table().erase((unsigned int)mxGetScalar(prhs[1]));
        clode_destroy(s);
        mexUnlock();
        plhs[0] = mxCreateLogicalScalar(table().empty());
        return;
    }

    clode_sim *s = instance(nrhs, prhs);
    if (action == "setproblemInfo" || action == "setprobleminfo") {
        need(nrhs, 3, action);
        Problem p;
        problem(prhs[2], p);
        check(clode_set_problem_info(s, &p.pi));
    } else if (action == "setstepper") {
        need(nrhs, 3, action);
        check(clode_set_stepper(s, str(prhs[2]).c_str()));
    } else if (action == "setprecision") {
        need(nrhs, 3, action);
        check(clode_set_precision(s, scalar(prhs[2], "precision") != 0 ? CLODE_PREC_SINGLE : CLODE_PREC_DOUBLE));
    } else if (action == "setopencl") {
        need(nrhs, 4, action);
        check(clode_set_device(s, (int)scalar(prhs[3], "deviceID")));
    } else if (action == "buildcl") {
        check(clode_build(s));
    } else if (action == "initialize") {
        // (h, tspan, x0, pars, sp)
        need(nrhs, 6, action);
        std::vector<double> t = vec(prhs[2]), x0 = vec(prhs[3]), pars = vec(prhs[4]);
        check(clode_set_tspan(s, t.data(), t.size()));
        check(clode_set_problem_data(s, x0.data(), x0.size(), pars.data(), pars.size()));
        clode_solver_params sp = solver(prhs[5]);
        check(clode_set_solver_params(s, &sp));
    } else if (action == "setnpts") {
        need(nrhs, 3, action);
        int nv = 0, np = 0, na = 0, nw = 0;
        check(clode_get_problem_info(s, &nv, &np, &na, &nw));
        const size_t n = (size_t)scalar(prhs[2], "nPts");
        std::vector<double> x0(n * nv, 0.0), pars(n * np, 0.0);
        check(clode_set_problem_data(s, x0.data(), x0.size(), pars.data(), pars.size()));
    } else if (action == "setproblemdata") {
        need(nrhs, 4, action);
        std::vector<double> x0 = vec(prhs[2]), pars = vec(prhs[3]);
        check(clode_set_problem_data(s, x0.data(), x0.size(), pars.data(), pars.size()));
    } else if (action == "settspan") {
        need(nrhs, 3, action);
        std::vector<double> t = vec(prhs[2]);
        check(clode_set_tspan(s, t.data(), t.size()));
    } else if (action == "setx0") {
        need(nrhs, 3, action);
        std::vector<double> x0 = vec(prhs[2]);
        check(clode_set_x0(s, x0.data(), x0.size()));
    } else if (action == "setpars") {
        need(nrhs, 3, action);
        std::vector<double> p = vec(prhs[2]);
        check(clode_set_pars(s, p.data(), p.size()));
    } else if (action == "setsolverpars") {
        need(nrhs, 3, action);
        clode_solver_params sp = solver(prhs[2]);
        check(clode_set_solver_params(s, &sp));
    } else if (action == "seedrng") {
        if (nrhs == 2)
            check(clode_seed_rng(s));
        else
            check(clode_seed_rng_int(s, (int)scalar(prhs[2], "seed")));
    } else if (action == "transient") {
        check(clode_transient(s));
    } else if (action == "shifttspan") {
        check(clode_shift_tspan(s));
    } else if (action == "shiftx0") {
        check(clode_shift_x0(s));
    } else if (action == "gettspan") {
        std::vector<double> t(2);
        check(clode_get_tspan(s, t.data()));
        plhs[0] = row(t);
    } else if (action == "getx0" || action == "getxf") {
        int nv = 0, np = 0, na = 0, nw = 0;
        check(clode_get_problem_info(s, &nv, &np, &na, &nw));
        std::vector<double> v(members(s) * nv);
        check(action == "getx0" ? clode_get_x0(s, v.data(), v.size()) : clode_get_xf(s, v.data(), v.size()));
        plhs[0] = row(v);
    } else if (action == "getsteppernames") {
        plhs[0] = name_cell([&](char *b, size_t c, size_t *n) { return clode_get_available_steppers(s, b, c, n); });
    } else if (action == "getprogramstring") {
        const std::string buf =
            get_string([&](char *b, size_t c, size_t *n) { return clode_get_program_string(s, b, c, n); });
        plhs[0] = mxCreateCellMatrix(1, 1);
        mxSetCell(plhs[0], 0, mxCreateString(buf.c_str()));
    } else if (action == "printstatus") {
        check(clode_print_status(s));
#if CLODE_MEX_KIND == 1
    } else if (action == "setobserverpars") {
        need(nrhs, 3, action);
        clode_observer_params op = observer(prhs[2]);
        check(clode_set_observer_params(s, &op));
    } else if (action == "setobserver") {
        need(nrhs, 3, action);
        check(clode_set_observer(s, str(prhs[2]).c_str()));
    } else if (action == "initobserver") {
        check(clode_initialize_observer(s));
    } else if (action == "features") {
        if (nrhs == 2) {
            check(clode_features(s, -1));
        } else {
            const double v = scalar(prhs[2], "initialization flag");
            if (v != 0 && v != 1)
                fail("Argument must be 0 or 1 for features with observer data initialization control");
            check(clode_features(s, (int)v));
        }
    } else if (action == "getnfeatures") {
        int n = 0;
        check(clode_get_n_features(s, &n));
        plhs[0] = mxCreateDoubleScalar((double)n);
    } else if (action == "getf") {
        int nf = 0;
        check(clode_get_n_features(s, &nf));
        std::vector<double> F((size_t)nf * members(s));
        check(clode_get_f(s, F.data(), F.size()));
        plhs[0] = column(F);
    } else if (action == "getfeaturenames") {
        plhs[0] = name_cell([&](char *b, size_t c, size_t *n) { return clode_get_feature_names(s, b, c, n); });
    } else if (action == "getobservernames") {
        plhs[0] = name_cell([&](char *b, size_t c, size_t *n) { return clode_get_available_observers(s, b, c, n); });
#endif
#if CLODE_MEX_KIND == 2
    } else if (action == "trajectory") {
        check(clode_trajectory(s));
    } else if (action == "gett" || action == "getx" || action == "getdx" || action == "getaux") {
        size_t nt = 0, nx = 0, na = 0;
        check(clode_get_trajectory_sizes(s, &nt, &nx, &na));
        std::vector<double> v(action == "gett" ? nt : action == "getaux" ? na : nx);
        if (action == "gett")
            check(clode_get_t(s, v.data(), v.size()));
        else if (action == "getx")
            check(clode_get_x(s, v.data(), v.size()));
        else if (action == "getdx")
            check(clode_get_dx(s, v.data(), v.size()));
        else
            check(clode_get_aux(s, v.data(), v.size()));
        plhs[0] = column(v);
    } else if (action == "getnstored") {
        std::vector<int> ns(members(s));
        check(clode_get_n_stored(s, ns.data(), ns.size()));
        plhs[0] = column(std::vector<double>(ns.begin(), ns.end()));
#endif
    } else {
        fail("Unrecognized action: " + action);
    }
}
