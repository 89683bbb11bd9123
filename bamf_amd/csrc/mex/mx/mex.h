// A host implementation of the subset of MATLAB's mx/mex C API that the
// wrappers in bamf_amd/csrc/mex use, so that they compile and run outside
// MATLAB (tests drive them through the mxh_* entry points of mx.cpp).  Names
// and signatures are MATLAB's documented interface; the implementation
// (mx.cpp) is this project's: arrays are column-major with a class, a
// dimension vector and owned storage; a struct array owns one value per
// element and field; mexErrMsgIdAndTxt throws, and mxh_call turns the throw
// into an error return.
#pragma once
#include <cstddef>
#include <cstdint>

typedef size_t mwSize;
typedef size_t mwIndex;
typedef struct mxArray_tag mxArray;

typedef enum {
    mxUNKNOWN_CLASS = 0,
    mxCELL_CLASS,
    mxSTRUCT_CLASS,
    mxLOGICAL_CLASS,
    mxCHAR_CLASS,
    mxVOID_CLASS,
    mxDOUBLE_CLASS,
    mxSINGLE_CLASS,
    mxINT8_CLASS,
    mxUINT8_CLASS,
    mxINT16_CLASS,
    mxUINT16_CLASS,
    mxINT32_CLASS,
    mxUINT32_CLASS,
    mxINT64_CLASS,
    mxUINT64_CLASS,
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity flag);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity flag);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray* mxCreateDoubleScalar(double v);
mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names);
void mxDestroyArray(mxArray* a);

void* mxGetData(const mxArray* a);
double* mxGetPr(const mxArray* a);
double mxGetScalar(const mxArray* a);
int mxGetString(const mxArray* a, char* buf, mwSize buflen);
mwSize mxGetM(const mxArray* a);
mwSize mxGetN(const mxArray* a);
mwSize mxGetNumberOfDimensions(const mxArray* a);
const mwSize* mxGetDimensions(const mxArray* a);
size_t mxGetNumberOfElements(const mxArray* a);
mxClassID mxGetClassID(const mxArray* a);
bool mxIsSingle(const mxArray* a);
bool mxIsDouble(const mxArray* a);
bool mxIsInt32(const mxArray* a);
bool mxIsStruct(const mxArray* a);
bool mxIsChar(const mxArray* a);

mxArray* mxGetField(const mxArray* s, mwIndex i, const char* name);
void mxSetField(mxArray* s, mwIndex i, const char* name, mxArray* v);
void mxSetFieldByNumber(mxArray* s, mwIndex i, int k, mxArray* v);
int mxGetNumberOfFields(const mxArray* s);
const char* mxGetFieldNameByNumber(const mxArray* s, int k);
mxArray* mxGetFieldByNumber(const mxArray* s, mwIndex i, int k);

[[noreturn]] void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);

// the entry point every mex file defines
extern "C" void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
