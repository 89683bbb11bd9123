// Host implementation of the mx/mex subset declared in mex.h, plus the
// mxh_* C entry points that build arguments, call a mex file's mexFunction
// and read its outputs (bamf_amd/mexsim.py binds them).  Compiled into each
// wrapper's shared library beside the wrapper itself.
#include "mex.h"

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

struct mxArray_tag {
    mxClassID cls = mxUNKNOWN_CLASS;
    std::vector<mwSize> dims;
    std::vector<unsigned char> bytes;   // numeric, logical and char data (column-major)
    std::vector<std::string> fields;    // struct field names
    std::vector<mxArray*> values;       // struct values, element-major: values[i * nfields + k]
};

namespace {

struct MexError : std::exception {
    std::string id, msg;
    MexError(std::string i, std::string m) : id(std::move(i)), msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

thread_local std::string g_error;

size_t elem_size(mxClassID c) {
    switch (c) {
        case mxDOUBLE_CLASS:
        case mxINT64_CLASS:
        case mxUINT64_CLASS: return 8;
        case mxSINGLE_CLASS:
        case mxINT32_CLASS:
        case mxUINT32_CLASS: return 4;
        case mxINT16_CLASS:
        case mxUINT16_CLASS:
        case mxCHAR_CLASS: return 2;  // MATLAB chars are 16-bit
        case mxINT8_CLASS:
        case mxUINT8_CLASS:
        case mxLOGICAL_CLASS: return 1;
        default: return 0;
    }
}

size_t numel_of(const std::vector<mwSize>& d) {
    size_t n = 1;
    for (mwSize v : d) n *= v;
    return n;
}

// MATLAB keeps at least two dimensions and drops trailing singleton ones past the second
std::vector<mwSize> norm_dims(mwSize nd, const mwSize* dims) {
    std::vector<mwSize> d(dims, dims + nd);
    while (d.size() < 2) d.push_back(1);
    while (d.size() > 2 && d.back() == 1) d.pop_back();
    return d;
}

template <class T>
double as_double(const unsigned char* p) {
    T v;
    std::memcpy(&v, p, sizeof(T));
    return (double)v;
}

int field_index(const mxArray* s, const char* name) {
    for (size_t k = 0; k < s->fields.size(); k++)
        if (s->fields[k] == name) return (int)k;
    return -1;
}

}  // namespace

mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity flag) {
    if (flag != mxREAL || elem_size(cls) == 0) mexErrMsgIdAndTxt("mx:class", "only real numeric arrays");
    mxArray* a = new mxArray_tag;
    a->cls = cls;
    a->dims = norm_dims(ndim, dims);
    a->bytes.assign(numel_of(a->dims) * elem_size(cls), 0);
    return a;
}

mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID cls, mxComplexity flag) {
    const mwSize d[2] = {m, n};
    return mxCreateNumericArray(2, d, cls, flag);
}

mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag) {
    return mxCreateNumericMatrix(m, n, mxDOUBLE_CLASS, flag);
}

mxArray* mxCreateDoubleScalar(double v) {
    mxArray* a = mxCreateDoubleMatrix(1, 1, mxREAL);
    std::memcpy(a->bytes.data(), &v, 8);
    return a;
}

mxArray* mxCreateStructMatrix(mwSize m, mwSize n, int nfields, const char** names) {
    mxArray* a = new mxArray_tag;
    a->cls = mxSTRUCT_CLASS;
    a->dims = {m, n};
    for (int k = 0; k < nfields; k++) a->fields.emplace_back(names[k]);
    a->values.assign(m * n * (size_t)nfields, nullptr);
    return a;
}

void mxDestroyArray(mxArray* a) {
    if (!a) return;
    for (mxArray* v : a->values) mxDestroyArray(v);
    delete a;
}

void* mxGetData(const mxArray* a) { return a->bytes.empty() ? nullptr : (void*)a->bytes.data(); }

double* mxGetPr(const mxArray* a) {
    if (a->cls != mxDOUBLE_CLASS) mexErrMsgIdAndTxt("mx:class", "mxGetPr of a non-double array");
    return (double*)mxGetData(a);
}

double mxGetScalar(const mxArray* a) {
    if (a->cls == mxSTRUCT_CLASS || a->bytes.empty()) return 0.0;
    const unsigned char* p = a->bytes.data();
    switch (a->cls) {
        case mxDOUBLE_CLASS: return as_double<double>(p);
        case mxSINGLE_CLASS: return as_double<float>(p);
        case mxINT64_CLASS: return as_double<int64_t>(p);
        case mxUINT64_CLASS: return as_double<uint64_t>(p);
        case mxINT32_CLASS: return as_double<int32_t>(p);
        case mxUINT32_CLASS: return as_double<uint32_t>(p);
        case mxINT16_CLASS: return as_double<int16_t>(p);
        case mxUINT16_CLASS:
        case mxCHAR_CLASS: return as_double<uint16_t>(p);
        case mxINT8_CLASS: return as_double<int8_t>(p);
        default: return as_double<uint8_t>(p);
    }
}

int mxGetString(const mxArray* a, char* buf, mwSize buflen) {
    if (a->cls != mxCHAR_CLASS || buflen == 0) return 1;
    const size_t n = numel_of(a->dims);
    const uint16_t* c = (const uint16_t*)a->bytes.data();
    size_t k = 0;
    for (; k < n && k + 1 < buflen; k++) buf[k] = (char)c[k];
    buf[k] = 0;
    return k < n ? 1 : 0;  // 1: truncated
}

mwSize mxGetM(const mxArray* a) { return a->dims[0]; }
mwSize mxGetN(const mxArray* a) { return numel_of(a->dims) / (a->dims[0] ? a->dims[0] : 1); }
mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->dims.size(); }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims.data(); }
size_t mxGetNumberOfElements(const mxArray* a) { return numel_of(a->dims); }
mxClassID mxGetClassID(const mxArray* a) { return a->cls; }
bool mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
bool mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsInt32(const mxArray* a) { return a->cls == mxINT32_CLASS; }
bool mxIsStruct(const mxArray* a) { return a->cls == mxSTRUCT_CLASS; }
bool mxIsChar(const mxArray* a) { return a->cls == mxCHAR_CLASS; }

mxArray* mxGetField(const mxArray* s, mwIndex i, const char* name) {
    if (s->cls != mxSTRUCT_CLASS) return nullptr;
    const int k = field_index(s, name);
    if (k < 0 || i >= numel_of(s->dims)) return nullptr;
    return s->values[i * s->fields.size() + (size_t)k];
}

void mxSetFieldByNumber(mxArray* s, mwIndex i, int k, mxArray* v) {
    if (s->cls != mxSTRUCT_CLASS || k < 0 || (size_t)k >= s->fields.size() || i >= numel_of(s->dims))
        mexErrMsgIdAndTxt("mx:field", "field %d of element %zu out of range", k, (size_t)i);
    mxArray*& slot = s->values[i * s->fields.size() + (size_t)k];
    if (slot != v) mxDestroyArray(slot);
    slot = v;
}

void mxSetField(mxArray* s, mwIndex i, const char* name, mxArray* v) {
    const int k = field_index(s, name);
    if (k < 0) mexErrMsgIdAndTxt("mx:field", "no field %s", name);
    mxSetFieldByNumber(s, i, k, v);
}

int mxGetNumberOfFields(const mxArray* s) { return (int)s->fields.size(); }
const char* mxGetFieldNameByNumber(const mxArray* s, int k) {
    return k >= 0 && (size_t)k < s->fields.size() ? s->fields[(size_t)k].c_str() : nullptr;
}
mxArray* mxGetFieldByNumber(const mxArray* s, mwIndex i, int k) {
    if (s->cls != mxSTRUCT_CLASS || k < 0 || (size_t)k >= s->fields.size() || i >= numel_of(s->dims)) return nullptr;
    return s->values[i * s->fields.size() + (size_t)k];
}

void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw MexError(id ? id : "", buf);
}

// ---------------------------------------------------------------- host side
extern "C" {

mxArray* mxh_numeric(int cls, int ndim, const size_t* dims, const void* data) {
    try {
        mxArray* a = mxCreateNumericArray((mwSize)ndim, dims, (mxClassID)cls, mxREAL);
        if (data && !a->bytes.empty()) std::memcpy(a->bytes.data(), data, a->bytes.size());
        return a;
    } catch (const MexError& e) {
        g_error = e.msg;
        return nullptr;
    }
}

mxArray* mxh_char(const char* s) {
    mxArray* a = new mxArray_tag;
    a->cls = mxCHAR_CLASS;
    const size_t n = std::strlen(s);
    a->dims = {1, n};
    a->bytes.resize(2 * n);
    uint16_t* c = (uint16_t*)a->bytes.data();
    for (size_t k = 0; k < n; k++) c[k] = (uint16_t)(unsigned char)s[k];
    return a;
}

mxArray* mxh_struct(size_t nelem, int nfields, const char** names) {
    return mxCreateStructMatrix(nelem, 1, nfields, names);
}

int mxh_set_field(mxArray* s, size_t i, int k, mxArray* v) {
    try {
        mxSetFieldByNumber(s, i, k, v);
        return 0;
    } catch (const MexError& e) {
        g_error = e.msg;
        return 1;
    }
}

// Calls the mex file's entry; 0 on return, 1 when it raised an error (message
// in mxh_last_error; outputs it created before the error are destroyed).
// plhs holds nslots >= max(nlhs, 1) entries.
int mxh_call(int nlhs, mxArray** plhs, int nslots, int nrhs, const mxArray** prhs) {
    const int nout = nslots;
    for (int k = 0; k < nout; k++) plhs[k] = nullptr;
    try {
        mexFunction(nlhs, plhs, nrhs, prhs);
        return 0;
    } catch (const MexError& e) {
        g_error = e.id + ": " + e.msg;
    } catch (const std::exception& e) {
        g_error = std::string("exception: ") + e.what();
    }
    for (int k = 0; k < nout; k++) {
        mxDestroyArray(plhs[k]);
        plhs[k] = nullptr;
    }
    return 1;
}

const char* mxh_last_error() { return g_error.c_str(); }
int mxh_class(const mxArray* a) { return (int)a->cls; }
int mxh_ndims(const mxArray* a) { return (int)a->dims.size(); }
const size_t* mxh_dims(const mxArray* a) { return a->dims.data(); }
const void* mxh_data(const mxArray* a) { return a->bytes.empty() ? nullptr : a->bytes.data(); }
size_t mxh_nbytes(const mxArray* a) { return a->bytes.size(); }
int mxh_nfields(const mxArray* a) { return (int)a->fields.size(); }
const char* mxh_field_name(const mxArray* a, int k) { return mxGetFieldNameByNumber(a, k); }
const mxArray* mxh_field_at(const mxArray* a, size_t i, int k) { return mxGetFieldByNumber(a, i, k); }
void mxh_destroy(mxArray* a) { mxDestroyArray(a); }

}  // extern "C"
