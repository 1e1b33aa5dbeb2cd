/* mex.h — minimal stand-in of MATLAB's interleaved-complex MEX API (the
 * subset used by ustb_amd/mex/das_hip.cpp), for compiling and exercising that
 * gateway in tests without MATLAB. Not MATLAB's header; behaviour follows the
 * documented R2018a API. mexErrMsgIdAndTxt throws a C++ exception carrying the
 * identifier (MATLAB longjmps back to the prompt instead). */
#ifndef USTBX_TEST_MEX_SHIM_H
#define USTBX_TEST_MEX_SHIM_H
#include <stddef.h>
#include <stdint.h>

typedef size_t mwSize;
typedef enum { mxDOUBLE_CLASS = 6, mxSINGLE_CLASS = 7, mxINT32_CLASS = 12 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;
typedef struct {
  float real, imag;
} mxComplexSingle;
typedef struct mxArray_tag mxArray;

mwSize mxGetNumberOfDimensions(const mxArray*);
const mwSize* mxGetDimensions(const mxArray*);
size_t mxGetNumberOfElements(const mxArray*);
bool mxIsDouble(const mxArray*);
bool mxIsSingle(const mxArray*);
bool mxIsInt32(const mxArray*);
bool mxIsComplex(const mxArray*);
float* mxGetSingles(const mxArray*);
mxComplexSingle* mxGetComplexSingles(const mxArray*);
int32_t* mxGetInt32s(const mxArray*);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity cplx);
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...);
void mexErrMsgTxt(const char* msg);

#endif
