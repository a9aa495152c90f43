/* Stand-in for MATLAB's mex.h (MATLAB is not installed here), declaring the
 * subset of the MEX/MX API that clode_amd/matlab/clode_mex.cpp uses, with
 * MATLAB's signatures.  tests/mex_stub/fake_mex.cpp implements it so the
 * gateways can be compiled and driven from the tests. */
#ifndef CLODE_TEST_MEX_STUB_H
#define CLODE_TEST_MEX_STUB_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

void mexErrMsgTxt(const char *msg);
int mexPrintf(const char *fmt, ...);
void mexLock(void);
void mexUnlock(void);

int mxIsChar(const mxArray *a);
int mxIsDouble(const mxArray *a);
int mxIsStruct(const mxArray *a);
int mxIsCell(const mxArray *a);
char *mxArrayToString(const mxArray *a);
void mxFree(void *p);
size_t mxGetNumberOfElements(const mxArray *a);
double mxGetScalar(const mxArray *a);
double *mxGetPr(const mxArray *a);
mxArray *mxGetField(const mxArray *s, mwIndex i, const char *name);
mxArray *mxGetCell(const mxArray *c, mwIndex i);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c);
mxArray *mxCreateDoubleScalar(double v);
mxArray *mxCreateLogicalScalar(int v);
mxArray *mxCreateCellMatrix(mwSize m, mwSize n);
void mxSetCell(mxArray *c, mwIndex i, mxArray *v);
mxArray *mxCreateString(const char *s);

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#ifdef __cplusplus
}
#endif

#endif
