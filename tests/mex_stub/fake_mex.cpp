// A minimal in-process implementation of the MEX/MX API subset declared in
// tests/mex_stub/mex.h, plus a C interface (fake_*) through which the tests
// build arguments, call a gateway's mexFunction and read its outputs.
// mexErrMsgTxt throws; fake_call catches it and reports the message, which is
// how MATLAB surfaces a gateway error to the caller.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
    enum Kind { DOUBLE, CHAR, CELL, STRUCT, LOGICAL } kind = DOUBLE;
    size_t m = 0, n = 0;
    std::vector<double> data;
    std::string str;
    std::vector<mxArray *> cells;
    std::map<std::string, mxArray *> fields;
};

namespace {
struct MexError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
int g_lock = 0;
std::string g_err;
}  // namespace

extern "C" {

void mexErrMsgTxt(const char *msg) { throw MexError(msg ? msg : ""); }
int mexPrintf(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = vprintf(fmt, ap);
    va_end(ap);
    return r;
}
void mexLock(void) { ++g_lock; }
void mexUnlock(void) { --g_lock; }

int mxIsChar(const mxArray *a) { return a && a->kind == mxArray::CHAR; }
int mxIsDouble(const mxArray *a) { return a && a->kind == mxArray::DOUBLE; }
int mxIsStruct(const mxArray *a) { return a && a->kind == mxArray::STRUCT; }
int mxIsCell(const mxArray *a) { return a && a->kind == mxArray::CELL; }
char *mxArrayToString(const mxArray *a)
{
    if (!mxIsChar(a))
        return nullptr;
    char *c = (char *)std::malloc(a->str.size() + 1);
    std::memcpy(c, a->str.c_str(), a->str.size() + 1);
    return c;
}
void mxFree(void *p) { std::free(p); }
size_t mxGetNumberOfElements(const mxArray *a)
{
    if (!a)
        return 0;
    if (a->kind == mxArray::CHAR)
        return a->str.size();
    if (a->kind == mxArray::CELL)
        return a->cells.size();
    if (a->kind == mxArray::STRUCT)
        return 1;
    return a->data.size();
}
double mxGetScalar(const mxArray *a) { return a && !a->data.empty() ? a->data[0] : 0.0; }
double *mxGetPr(const mxArray *a) { return a ? const_cast<double *>(a->data.data()) : nullptr; }
mxArray *mxGetField(const mxArray *s, mwIndex, const char *name)
{
    if (!mxIsStruct(s))
        return nullptr;
    auto it = s->fields.find(name);
    return it == s->fields.end() ? nullptr : it->second;
}
mxArray *mxGetCell(const mxArray *c, mwIndex i) { return mxIsCell(c) && i < c->cells.size() ? c->cells[i] : nullptr; }
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity)
{
    auto *a = new mxArray;
    a->m = m;
    a->n = n;
    a->data.assign(m * n, 0.0);
    return a;
}
mxArray *mxCreateDoubleScalar(double v)
{
    mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
    a->data[0] = v;
    return a;
}
mxArray *mxCreateLogicalScalar(int v)
{
    mxArray *a = mxCreateDoubleScalar(v ? 1.0 : 0.0);
    a->kind = mxArray::LOGICAL;
    return a;
}
mxArray *mxCreateCellMatrix(mwSize m, mwSize n)
{
    auto *a = new mxArray;
    a->kind = mxArray::CELL;
    a->m = m;
    a->n = n;
    a->cells.assign(m * n, nullptr);
    return a;
}
void mxSetCell(mxArray *c, mwIndex i, mxArray *v)
{
    if (mxIsCell(c) && i < c->cells.size())
        c->cells[i] = v;
}
mxArray *mxCreateString(const char *s)
{
    auto *a = new mxArray;
    a->kind = mxArray::CHAR;
    a->str = s ? s : "";
    a->m = 1;
    a->n = a->str.size();
    return a;
}

// ---- test driver interface
mxArray *fake_double(const double *v, size_t m, size_t n)
{
    mxArray *a = mxCreateDoubleMatrix(m, n, mxREAL);
    if (v)
        std::memcpy(a->data.data(), v, m * n * sizeof(double));
    return a;
}
mxArray *fake_string(const char *s) { return mxCreateString(s); }
mxArray *fake_cell(size_t n) { return mxCreateCellMatrix(1, n); }
void fake_cell_set(mxArray *c, size_t i, mxArray *v) { mxSetCell(c, i, v); }
mxArray *fake_struct(void)
{
    auto *a = new mxArray;
    a->kind = mxArray::STRUCT;
    a->m = a->n = 1;
    return a;
}
void fake_struct_set(mxArray *s, const char *name, mxArray *v) { s->fields[name] = v; }
int fake_call(int nlhs, mxArray **plhs, int nrhs, mxArray **prhs)
{
    try {
        mexFunction(nlhs, plhs, nrhs, const_cast<const mxArray **>(prhs));
    } catch (const std::exception &e) {
        g_err = e.what();
        return 1;
    }
    return 0;
}
const char *fake_last_error(void) { return g_err.c_str(); }
int fake_lock_count(void) { return g_lock; }
size_t fake_numel(const mxArray *a) { return mxGetNumberOfElements(a); }
const double *fake_data(const mxArray *a) { return a ? a->data.data() : nullptr; }
int fake_is_cell(const mxArray *a) { return mxIsCell(a); }
mxArray *fake_cell_get(const mxArray *a, size_t i) { return mxGetCell(a, i); }
const char *fake_str(const mxArray *a) { return a ? a->str.c_str() : ""; }
}
