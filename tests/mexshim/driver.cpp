// driver.cpp — implements the mex.h stand-in and drives the das_hip gateway.
//   driver errors : argument-error scenarios (no GPU needed); prints
//                   "<scenario> <identifier>" per line
//   driver run    : one call per sum mode on seeded data, compared with the
//                   das_c restatement of the oracle (GPU needed); prints
//                   "dim <d> rel <max|gpu-oracle|/max|oracle|>"
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "mex.h"

struct mxArray_tag {
  mxClassID cls;
  bool cplx;
  std::vector<mwSize> dims;
  std::vector<float> f;        // single (interleaved if complex)
  std::vector<double> d;       // double
  std::vector<int32_t> i;      // int32
};

struct MexError : std::runtime_error {
  std::string id;
  MexError(const char* i, const std::string& m) : std::runtime_error(m), id(i) {}
};

mwSize mxGetNumberOfDimensions(const mxArray* a) { return a->dims.size(); }
const mwSize* mxGetDimensions(const mxArray* a) { return a->dims.data(); }
size_t mxGetNumberOfElements(const mxArray* a) {
  size_t n = 1;
  for (mwSize v : a->dims) n *= v;
  return n;
}
bool mxIsDouble(const mxArray* a) { return a->cls == mxDOUBLE_CLASS; }
bool mxIsSingle(const mxArray* a) { return a->cls == mxSINGLE_CLASS; }
bool mxIsInt32(const mxArray* a) { return a->cls == mxINT32_CLASS; }
bool mxIsComplex(const mxArray* a) { return a->cplx; }
float* mxGetSingles(const mxArray* a) { return a->cls == mxSINGLE_CLASS && !a->cplx ? (float*)a->f.data() : nullptr; }
mxComplexSingle* mxGetComplexSingles(const mxArray* a) {
  return a->cls == mxSINGLE_CLASS && a->cplx ? (mxComplexSingle*)a->f.data() : nullptr;
}
int32_t* mxGetInt32s(const mxArray* a) { return a->cls == mxINT32_CLASS ? (int32_t*)a->i.data() : nullptr; }
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID cls, mxComplexity cplx) {
  mxArray* a = new mxArray_tag{cls, cplx == mxCOMPLEX, std::vector<mwSize>(dims, dims + ndim), {}, {}, {}};
  const size_t n = mxGetNumberOfElements(a);
  if (cls == mxSINGLE_CLASS) a->f.assign(n * (a->cplx ? 2 : 1), 0.0f);
  if (cls == mxDOUBLE_CLASS) a->d.assign(n, 0.0);
  if (cls == mxINT32_CLASS) a->i.assign(n, 0);
  return a;
}
void mexErrMsgIdAndTxt(const char* id, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw MexError(id, buf);
}
void mexErrMsgTxt(const char* msg) { throw MexError("", msg); }

void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
extern "C" int oracle_das_c(const float* ch, int L, int N, int W, int F, int P, float Fs, float t0, float fd,
                            int sum_mode, const float* tx_apo, const float* rx_apo, const float* tx_delay,
                            const float* rx_delay, float* out, int nthreads);

static mxArray* single(std::vector<mwSize> dims, bool cplx = false) {
  return mxCreateNumericArray(dims.size(), dims.data(), mxSINGLE_CLASS, cplx ? mxCOMPLEX : mxREAL);
}
static mxArray* scalar(float v) {
  mxArray* a = single({1, 1});
  a->f[0] = v;
  return a;
}
static mxArray* i32(int v) {
  mwSize d[2] = {1, 1};
  mxArray* a = mxCreateNumericArray(2, d, mxINT32_CLASS, mxREAL);
  a->i[0] = v;
  return a;
}

struct Args {
  std::vector<mxArray*> a;
  Args(int L, int N, int W, int F, int P, float fd, int dim, unsigned seed) {
    std::mt19937 g(seed);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    mxArray* ch = single({(mwSize)L, (mwSize)N, (mwSize)W, (mwSize)F}, true);
    for (auto& v : ch->f) v = u(g) * 2 - 1;
    mxArray* ta = single({(mwSize)P, (mwSize)W});
    mxArray* ra = single({(mwSize)P, (mwSize)N});
    mxArray* td = single({(mwSize)P, (mwSize)W});
    mxArray* rd = single({(mwSize)P, (mwSize)N});
    for (auto& v : ta->f) v = u(g) < 0.2f ? 0.0f : u(g);
    for (auto& v : ra->f) v = u(g) < 0.2f ? 0.0f : u(g);
    for (auto& v : td->f) v = u(g) * L * 0.45f;
    for (auto& v : rd->f) v = u(g) * L * 0.45f;
    a = {ch, scalar(1.0f), scalar(0.0f), ta, ra, td, rd, scalar(fd), i32(dim)};
  }
};

static const char* call(std::vector<mxArray*> prhs, int nlhs, mxArray** out) {
  static std::string id;
  try {
    mxArray* plhs[1] = {nullptr};
    mexFunction(nlhs, plhs, (int)prhs.size(), (const mxArray**)prhs.data());
    if (out) *out = plhs[0];
    return "ok";
  } catch (const MexError& e) {
    id = e.id.empty() ? "(no-identifier)" : e.id;
    return id.c_str();
  }
}

int main(int argc, char** argv) {
  const std::string mode = argc > 1 ? argv[1] : "errors";
  if (mode == "errors") {
    Args A(32, 4, 2, 1, 10, 0.1f, 3, 1);
    auto v = A.a;
    printf("nrhs8 %s\n", call(std::vector<mxArray*>(v.begin(), v.begin() + 8), 1, nullptr));
    printf("nlhs2 %s\n", call(v, 2, nullptr));
    auto w = v;
    w[8] = i32(7);
    printf("summode %s\n", call(w, 1, nullptr));
    w = v;
    w[5] = single({10, 3});
    printf("txdelay_waves %s\n", call(w, 1, nullptr));
    w = v;
    w[6] = single({9, 4});
    printf("rxdelay_pixels %s\n", call(w, 1, nullptr));
    w = v;
    w[1] = mxCreateNumericArray(2, std::vector<mwSize>{1, 1}.data(), mxDOUBLE_CLASS, mxREAL);
    printf("fs_double %s\n", call(w, 1, nullptr));
    w = v;
    w[0] = single({32, 4, 2, 1}, false);
    printf("real_with_fd %s\n", call(w, 1, nullptr));
    w = v;
    w[7] = single({1, 2});
    printf("fd_vector %s\n", call(w, 1, nullptr));
    return 0;
  }
  // run: all four modes vs the oracle
  const int L = 96, N = 12, W = 5, F = 2, P = 500;
  for (int dim = 0; dim < 4; ++dim) {
    for (int iq = 0; iq < 2; ++iq) {
      const float fd = iq ? 0.05f : 0.0f;
      Args A(L, N, W, F, P, fd, dim, 7 + dim);
      if (!iq) {  // RF: real channel data is promoted by the gateway
        mxArray* re = single({(mwSize)L, (mwSize)N, (mwSize)W, (mwSize)F});
        for (size_t k = 0; k < re->f.size(); ++k) re->f[k] = A.a[0]->f[2 * k];
        A.a[0] = re;
      }
      mxArray* out = nullptr;
      const char* r = call(A.a, 1, &out);
      if (strcmp(r, "ok") != 0) {
        printf("dim %d iq %d error %s\n", dim, iq, r);
        return 1;
      }
      std::vector<float> ch(2 * (size_t)L * N * W * F, 0.0f);
      for (size_t k = 0; k < ch.size() / 2; ++k) {
        ch[2 * k] = iq ? A.a[0]->f[2 * k] : A.a[0]->f[k];
        ch[2 * k + 1] = iq ? A.a[0]->f[2 * k + 1] : 0.0f;
      }
      std::vector<float> ref(out->f.size(), 0.0f);
      oracle_das_c(ch.data(), L, N, W, F, P, 1.0f, 0.0f, fd, dim, A.a[3]->f.data(), A.a[4]->f.data(),
                   A.a[5]->f.data(), A.a[6]->f.data(), ref.data(), 4);
      double md = 0, mr = 0;
      for (size_t k = 0; k < ref.size(); ++k) {
        md = std::fmax(md, std::fabs((double)out->f[k] - ref[k]));
        mr = std::fmax(mr, std::fabs((double)ref[k]));
      }
      printf("dim %d iq %d dims %zu %zu %zu %zu rel %.3e\n", dim, iq, out->dims[0], out->dims[1], out->dims[2],
             out->dims[3], md / mr);
    }
  }
  return 0;
}
