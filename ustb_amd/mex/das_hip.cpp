// das_hip.cpp — MATLAB MEX gateway of the MI355X beamformer.
//
// Call signature of the reference MEX files (+midprocess/das.m:251-271):
//   bf = das_hip(ch, fs, t0, tx_apo, rx_apo, tx_delay, rx_delay, fd, int32(dim))
//   bf = das_hip(..., int32(dim), int32(gpu_id))
// ch [time, channel, wave, frame] single (complex, or real with fd == 0);
// fs, t0, fd single scalars; tx_apo/tx_delay [pixel, wave] single;
// rx_apo/rx_delay [pixel, channel] single; dim: 0 none, 1 receive,
// 2 transmit, 3 both (all four, unlike das_cuda's both-only,
// das_cuda.cu:321-324). Output: complex single [pixel, N', W', frame].
//
// Argument checks follow das_c (+mex/source/das_c.cpp:53-252) through
// ustbx_das_check_mex_args; errors are raised as das_c raises them, with its
// "Toolbox:SRP_SRC:*" identifier or (das_c's mexErrMsgTxt cases) none. The 10th argument is das_cuda's gpu_id
// (das_cuda.cu:161-162); das_c's (dead) verbose flag in that position is
// not supported.
//
// Build (interleaved-complex API, as build_mex.m:45 uses):
//   mex -R2018a das_hip.cpp -I<repo>/include -L<repo>/ustb_amd -lustbx
#include <stdint.h>
#include <string.h>

#include <vector>

#include "mex.h"
#include "ustbx_das.h"

// das_c raises some errors with an identifier and the others through
// mexErrMsgTxt (ustbx_status_id gives "" for those); so does the gateway.
static void fail(int code, const char* msg) {
  const char* id = ustbx_status_id(code);
  if (id[0])
    mexErrMsgIdAndTxt(id, "%s", msg);
  else
    mexErrMsgTxt(msg);
}

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]) {
  if (nrhs < 9 || nrhs > 10 || nlhs > 1) {
    // let the shared checker produce the reference message
    int64_t z[4] = {0, 0, 0, 0};
    int32_t nd[4] = {2, 2, 2, 2}, dbl[4] = {0, 0, 0, 0}, sdbl[3] = {0, 0, 0};
    int64_t md[8] = {0, 0, 0, 0, 0, 0, 0, 0}, sn[3] = {1, 1, 1};
    ustbx_das_desc d;
    ustbx_status st;
    const int rc = ustbx_das_check_mex_args(nrhs, nlhs, 2, z, 0, nd, md, dbl, sn, sdbl, 3, 1, 0.0f, &d, &st);
    fail(rc ? rc : USTBX_E_NRHS, st.message);
    return;
  }
  const mxArray* ch = prhs[0];
  int64_t ch_dims[4] = {1, 1, 1, 1};
  const int ch_ndim = (int)mxGetNumberOfDimensions(ch);
  const mwSize* cd = mxGetDimensions(ch);
  for (int k = 0; k < ch_ndim && k < 4; ++k) ch_dims[k] = (int64_t)cd[k];
  int32_t mat_ndim[4], mat_dbl[4];
  int64_t mat_dims[8];
  for (int k = 0; k < 4; ++k) {  // prhs[3..6]: tx_apo, rx_apo, tx_delay, rx_delay
    const mxArray* m = prhs[3 + k];
    mat_ndim[k] = (int32_t)mxGetNumberOfDimensions(m);
    const mwSize* md = mxGetDimensions(m);
    mat_dims[2 * k] = (int64_t)md[0];
    mat_dims[2 * k + 1] = mat_ndim[k] >= 2 ? (int64_t)md[1] : 1;
    mat_dbl[k] = mxIsDouble(m) || !mxIsSingle(m) ? 1 : 0;
  }
  const mxArray* sc[3] = {prhs[1], prhs[2], prhs[7]};
  int64_t sn[3];
  int32_t sdbl[3];
  for (int k = 0; k < 3; ++k) {
    sn[k] = (int64_t)mxGetNumberOfElements(sc[k]);
    sdbl[k] = mxIsDouble(sc[k]) || !mxIsSingle(sc[k]) ? 1 : 0;
  }
  if (!mxIsInt32(prhs[8]) || mxGetNumberOfElements(prhs[8]) != 1)
    fail(USTBX_E_ARG, "The sum mode must be an int32 scalar");
  const int32_t dim = *mxGetInt32s(prhs[8]);
  const float fd = sdbl[2] ? 0.0f : *mxGetSingles(prhs[7]);
  ustbx_das_desc d;
  memset(&d, 0, sizeof(d));
  ustbx_status st;
  int rc = ustbx_das_check_mex_args(nrhs, nlhs, ch_ndim, ch_dims, mxIsDouble(ch) || !mxIsSingle(ch) ? 1 : 0,
                                    mat_ndim, mat_dims, mat_dbl, sn, sdbl, dim, mxIsComplex(ch) ? 1 : 0, fd,
                                    &d, &st);
  if (rc) fail(rc, st.message);
  int32_t gpu = 0;
  if (nrhs == 10) {
    if (!mxIsInt32(prhs[9]) || mxGetNumberOfElements(prhs[9]) != 1)
      fail(USTBX_E_ARG, "The gpu id must be an int32 scalar");
    gpu = *mxGetInt32s(prhs[9]);
  }
  d.fs = *mxGetSingles(prhs[1]);
  d.t0 = *mxGetSingles(prhs[2]);
  d.fd = fd;
  d.interp = USTBX_INTERP_LINEAR;
  d.edge = USTBX_EDGE_DASC;

  // real data (fd == 0 only, checked above): promote to complex with zero
  // imaginary part
  const size_t nch = (size_t)d.L * d.N * d.W * d.F;
  std::vector<float> promoted;
  const void* chp;
  if (mxIsComplex(ch)) {
    chp = mxGetComplexSingles(ch);
  } else {
    promoted.assign(2 * nch, 0.0f);
    const float* re = mxGetSingles(ch);
    for (size_t i = 0; i < nch; ++i) promoted[2 * i] = re[i];
    chp = promoted.data();
  }
  mwSize odims[4] = {(mwSize)d.P, (mwSize)((dim == 0 || dim == 2) ? d.N : 1),
                     (mwSize)((dim == 0 || dim == 1) ? d.W : 1), (mwSize)d.F};
  plhs[0] = mxCreateNumericArray(4, odims, mxSINGLE_CLASS, mxCOMPLEX);
  void* out = mxGetComplexSingles(plhs[0]);
  rc = ustbx_das_run_host(&d, chp, mxGetSingles(prhs[3]), mxGetSingles(prhs[4]), mxGetSingles(prhs[5]),
                          mxGetSingles(prhs[6]), out, gpu, &st);
  if (rc) fail(rc, st.message);
}
