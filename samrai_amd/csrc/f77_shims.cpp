// Link-compatible twins of the Fortran kernels the reference calls through
// SAMRAI_F77_FUNC (lower case, trailing underscore; every argument by
// address): solv/CellPoissonFACOps.C:62-543, solv/CartesianRobinBcHelper.C:32-73
// and the transfer operators (geom/CartesianCellDoubleWeightedAverage.C:37-70,
// geom/CartesianOutersideDoubleWeightedAverage.C:42-110,
// geom/CartesianCellDoubleConservativeLinearRefine.C:41-80,
// geom/CartesianCellDoubleLinearRefine.C, pdat/CellDoubleConstantRefine.C:40-60).
// A build of the reference that links
// this library instead of its Fortran objects runs the same calls on the GPU
// unchanged, provided its ArrayData storage is device memory (the reference
// allocates it through tbox::ResourceAllocator / Umpire, pdat/ArrayData.C:81-112).
//
// Each twin reads its scalar and index arguments on the host, runs the sfac_
// kernel of the same name on the null stream and waits for it, so it returns
// with its outputs written, as the Fortran routine does.  The red-black
// routines' maxres is a host double, as in the Fortran, read back from a
// per-device word.  The Fortran interface has no status: an error (invalid
// arguments, a failed launch) prints sfac_last_error() and aborts, as the
// reference's TBOX_ERROR does.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "samrai_f77.h"
#include "samrai_fac.h"

namespace {

void done(int rc, const char* name) {
  if (rc == 0 && hipStreamSynchronize(nullptr) == hipSuccess) return;
  std::fprintf(stderr, "samrai_amd %s: %s\n", name,
               rc ? sfac_last_error() : hipGetErrorString(hipGetLastError()));
  std::abort();
}

// One device double per GPU for the red-black routines' max |residual|
// (allocated on first use, held until the process ends).
double* maxres_word() {
  static std::mutex mu;
  static std::vector<double*> words;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> lk(mu);
  if ((int)words.size() <= dev) words.resize(dev + 1, nullptr);
  if (!words[dev] && hipMalloc(&words[dev], sizeof(double)) != hipSuccess) words[dev] = nullptr;
  return words[dev];
}

void read_maxres(const double* maxres, const char* name) {
  double* w = maxres_word();
  if (!w || hipMemcpy(const_cast<double*>(maxres), w, sizeof(double), hipMemcpyDeviceToHost) !=
                hipSuccess) {
    std::fprintf(stderr, "samrai_amd %s: maxres read-back failed\n", name);
    std::abort();
  }
}

double* maxres_or_abort(const char* name) {
  double* w = maxres_word();
  if (!w) {
    std::fprintf(stderr, "samrai_amd %s: no device word for maxres\n", name);
    std::abort();
  }
  return w;
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ 2-D

void compfluxvardc2d_(double* xflux, double* yflux, const int* fgi, const int* fgj,
                      const double* xdc, const double* ydc, const int* dcgi, const int* dcgj,
                      const double* soln, const int* sgi, const int* sgj, const int* ifirst,
                      const int* ilast, const int* jfirst, const int* jlast, const double* dx) {
  done(sfac_compfluxvardc2d(xflux, yflux, *fgi, *fgj, xdc, ydc, *dcgi, *dcgj, soln, *sgi, *sgj,
                            *ifirst, *ilast, *jfirst, *jlast, dx, nullptr),
       "compfluxvardc2d");
}

void compfluxcondc2d_(double* xflux, double* yflux, const int* fgi, const int* fgj,
                      const double* dc, const double* soln, const int* sgi, const int* sgj,
                      const int* ifirst, const int* ilast, const int* jfirst, const int* jlast,
                      const double* dx) {
  done(sfac_compfluxcondc2d(xflux, yflux, *fgi, *fgj, *dc, soln, *sgi, *sgj, *ifirst, *ilast,
                            *jfirst, *jlast, dx, nullptr),
       "compfluxcondc2d");
}

void rbgswithfluxmaxvardcvarsf2d_(const double* xflux, const double* yflux, const int* fgi,
                                  const int* fgj, const double* xdc, const double* ydc,
                                  const int* dcgi, const int* dcgj, const double* rhs,
                                  const int* rgi, const int* rgj, const double* sf,
                                  const int* sfgi, const int* sfgj, double* soln,
                                  const int* sgi, const int* sgj, const int* ifirst,
                                  const int* ilast, const int* jfirst, const int* jlast,
                                  const double* dx, const int* offset, const double* maxres) {
  const char* n = "rbgswithfluxmaxvardcvarsf2d";
  done(sfac_rbgswithfluxmax2d(xflux, yflux, *fgi, *fgj, xdc, ydc, *dcgi, *dcgj, 0.0, rhs, *rgi,
                              *rgj, sf, *sfgi, *sfgj, 0.0, soln, *sgi, *sgj, *ifirst, *ilast,
                              *jfirst, *jlast, dx, *offset, maxres_or_abort(n), nullptr),
       n);
  read_maxres(maxres, n);
}

void rbgswithfluxmaxcondcvarsf2d_(const double* xflux, const double* yflux, const int* fgi,
                                  const int* fgj, const double* dc, const double* rhs,
                                  const int* rgi, const int* rgj, const double* sf,
                                  const int* sfgi, const int* sfgj, double* soln,
                                  const int* sgi, const int* sgj, const int* ifirst,
                                  const int* ilast, const int* jfirst, const int* jlast,
                                  const double* dx, const int* offset, const double* maxres) {
  const char* n = "rbgswithfluxmaxcondcvarsf2d";
  done(sfac_rbgswithfluxmax2d(xflux, yflux, *fgi, *fgj, nullptr, nullptr, 0, 0, *dc, rhs, *rgi,
                              *rgj, sf, *sfgi, *sfgj, 0.0, soln, *sgi, *sgj, *ifirst, *ilast,
                              *jfirst, *jlast, dx, *offset, maxres_or_abort(n), nullptr),
       n);
  read_maxres(maxres, n);
}

void rbgswithfluxmaxvardcconsf2d_(const double* xflux, const double* yflux, const int* fgi,
                                  const int* fgj, const double* xdc, const double* ydc,
                                  const int* dcgi, const int* dcgj, const double* rhs,
                                  const int* rgi, const int* rgj, const double* sfc,
                                  double* soln, const int* sgi, const int* sgj,
                                  const int* ifirst, const int* ilast, const int* jfirst,
                                  const int* jlast, const double* dx, const int* offset,
                                  const double* maxres) {
  const char* n = "rbgswithfluxmaxvardcconsf2d";
  done(sfac_rbgswithfluxmax2d(xflux, yflux, *fgi, *fgj, xdc, ydc, *dcgi, *dcgj, 0.0, rhs, *rgi,
                              *rgj, nullptr, 0, 0, *sfc, soln, *sgi, *sgj, *ifirst, *ilast,
                              *jfirst, *jlast, dx, *offset, maxres_or_abort(n), nullptr),
       n);
  read_maxres(maxres, n);
}

void rbgswithfluxmaxcondcconsf2d_(const double* xflux, const double* yflux, const int* fgi,
                                  const int* fgj, const double* dc, const double* rhs,
                                  const int* rgi, const int* rgj, const double* sfc,
                                  double* soln, const int* sgi, const int* sgj,
                                  const int* ifirst, const int* ilast, const int* jfirst,
                                  const int* jlast, const double* dx, const int* offset,
                                  const double* maxres) {
  const char* n = "rbgswithfluxmaxcondcconsf2d";
  done(sfac_rbgswithfluxmax2d(xflux, yflux, *fgi, *fgj, nullptr, nullptr, 0, 0, *dc, rhs, *rgi,
                              *rgj, nullptr, 0, 0, *sfc, soln, *sgi, *sgj, *ifirst, *ilast,
                              *jfirst, *jlast, dx, *offset, maxres_or_abort(n), nullptr),
       n);
  read_maxres(maxres, n);
}

// residual takes the rhs ghost widths (the Fortran ignores residualgi/gj)
void compresvarsca2d_(const double* xflux, const double* yflux, const int* fgi, const int* fgj,
                      const double* rhs, const int* rgi, const int* rgj, double* residual,
                      const int* residualgi, const int* residualgj, const double* sf,
                      const int* sfgi, const int* sfgj, const double* soln, const int* sgi,
                      const int* sgj, const int* ifirst, const int* ilast, const int* jfirst,
                      const int* jlast, const double* dx) {
  (void)residualgi;
  (void)residualgj;
  done(sfac_compres2d(xflux, yflux, *fgi, *fgj, rhs, *rgi, *rgj, residual, sf, *sfgi, *sfgj, 0.0,
                      soln, *sgi, *sgj, *ifirst, *ilast, *jfirst, *jlast, dx, nullptr),
       "compresvarsca2d");
}

void compresconsca2d_(const double* xflux, const double* yflux, const int* fgi, const int* fgj,
                      const double* rhs, const int* rgi, const int* rgj, double* residual,
                      const int* residualgi, const int* residualgj, const double* sfc,
                      const double* soln, const int* sgi, const int* sgj, const int* ifirst,
                      const int* ilast, const int* jfirst, const int* jlast, const double* dx) {
  (void)residualgi;
  (void)residualgj;
  done(sfac_compres2d(xflux, yflux, *fgi, *fgj, rhs, *rgi, *rgj, residual, nullptr, 0, 0, *sfc,
                      soln, *sgi, *sgj, *ifirst, *ilast, *jfirst, *jlast, dx, nullptr),
       "compresconsca2d");
}

void ewingfixfluxvardc2d_(const double* xflux, const double* yflux, const int* fgi,
                          const int* fgj, const double* xdc, const double* ydc, const int* dcgi,
                          const int* dcgj, const double* soln, const int* sgi, const int* sgj,
                          const int* ifirst, const int* ilast, const int* jfirst,
                          const int* jlast, const int* location_index, const int* ratio,
                          const int* blower, const int* bupper, const double* dx) {
  done(sfac_ewingfixflux2d(const_cast<double*>(xflux), const_cast<double*>(yflux), *fgi, *fgj,
                           xdc, ydc, *dcgi, *dcgj, 0.0, soln, *sgi, *sgj, *ifirst, *ilast,
                           *jfirst, *jlast, *location_index, ratio, blower, bupper, dx, nullptr),
       "ewingfixfluxvardc2d");
}

void ewingfixfluxcondc2d_(const double* xflux, const double* yflux, const int* fgi,
                          const int* fgj, const double* dc, const double* soln, const int* sgi,
                          const int* sgj, const int* ifirst, const int* ilast,
                          const int* jfirst, const int* jlast, const int* location_index,
                          const int* ratio, const int* blower, const int* bupper,
                          const double* dx) {
  done(sfac_ewingfixflux2d(const_cast<double*>(xflux), const_cast<double*>(yflux), *fgi, *fgj,
                           nullptr, nullptr, 0, 0, *dc, soln, *sgi, *sgj, *ifirst, *ilast,
                           *jfirst, *jlast, *location_index, ratio, blower, bupper, dx, nullptr),
       "ewingfixfluxcondc2d");
}

// ------------------------------------------------------------------ 3-D

void compfluxvardc3d_(double* xflux, double* yflux, double* zflux, const int* fgi,
                      const int* fgj, const int* fgk, const double* xdc, const double* ydc,
                      const double* zdc, const int* dcgi, const int* dcgj, const int* dcgk,
                      const double* soln, const int* sgi, const int* sgj, const int* sgk,
                      const int* ifirst, const int* ilast, const int* jfirst, const int* jlast,
                      const int* kfirst, const int* klast, const double* dx) {
  done(sfac_compfluxvardc3d(xflux, yflux, zflux, *fgi, *fgj, *fgk, xdc, ydc, zdc, *dcgi, *dcgj,
                            *dcgk, soln, *sgi, *sgj, *sgk, *ifirst, *ilast, *jfirst, *jlast,
                            *kfirst, *klast, dx, nullptr),
       "compfluxvardc3d");
}

void compfluxcondc3d_(double* xflux, double* yflux, double* zflux, const int* fgi,
                      const int* fgj, const int* fgk, const double* dc, const double* soln,
                      const int* sgi, const int* sgj, const int* sgk, const int* ifirst,
                      const int* ilast, const int* jfirst, const int* jlast, const int* kfirst,
                      const int* klast, const double* dx) {
  done(sfac_compfluxcondc3d(xflux, yflux, zflux, *fgi, *fgj, *fgk, *dc, soln, *sgi, *sgj, *sgk,
                            *ifirst, *ilast, *jfirst, *jlast, *kfirst, *klast, dx, nullptr),
       "compfluxcondc3d");
}

void rbgswithfluxmaxvardcvarsf3d_(const double* xflux, const double* yflux, const double* zflux,
                                  const int* fgi, const int* fgj, const int* fgk,
                                  const double* xdc, const double* ydc, const double* zdc,
                                  const int* dcgi, const int* dcgj, const int* dcgk,
                                  const double* rhs, const int* rgi, const int* rgj,
                                  const int* rgk, const double* sf, const int* sfgi,
                                  const int* sfgj, const int* sfgk, double* soln,
                                  const int* sgi, const int* sgj, const int* sgk,
                                  const int* ifirst, const int* ilast, const int* jfirst,
                                  const int* jlast, const int* kfirst, const int* klast,
                                  const double* dx, const int* offset, const double* maxres) {
  const char* n = "rbgswithfluxmaxvardcvarsf3d";
  done(sfac_rbgswithfluxmax3d(xflux, yflux, zflux, *fgi, *fgj, *fgk, xdc, ydc, zdc, *dcgi, *dcgj,
                              *dcgk, 0.0, rhs, *rgi, *rgj, *rgk, sf, *sfgi, *sfgj, *sfgk, 0.0,
                              soln, *sgi, *sgj, *sgk, *ifirst, *ilast, *jfirst, *jlast, *kfirst,
                              *klast, dx, *offset, maxres_or_abort(n), nullptr),
       n);
  read_maxres(maxres, n);
}

void rbgswithfluxmaxcondcvarsf3d_(const double* xflux, const double* yflux, const double* zflux,
                                  const int* fgi, const int* fgj, const int* fgk,
                                  const double* dc, const double* rhs, const int* rgi,
                                  const int* rgj, const int* rgk, const double* sf,
                                  const int* sfgi, const int* sfgj, const int* sfgk,
                                  double* soln, const int* sgi, const int* sgj, const int* sgk,
                                  const int* ifirst, const int* ilast, const int* jfirst,
                                  const int* jlast, const int* kfirst, const int* klast,
                                  const double* dx, const int* offset, const double* maxres) {
  const char* n = "rbgswithfluxmaxcondcvarsf3d";
  done(sfac_rbgswithfluxmax3d(xflux, yflux, zflux, *fgi, *fgj, *fgk, nullptr, nullptr, nullptr,
                              0, 0, 0, *dc, rhs, *rgi, *rgj, *rgk, sf, *sfgi, *sfgj, *sfgk, 0.0,
                              soln, *sgi, *sgj, *sgk, *ifirst, *ilast, *jfirst, *jlast, *kfirst,
                              *klast, dx, *offset, maxres_or_abort(n), nullptr),
       n);
  read_maxres(maxres, n);
}

void rbgswithfluxmaxvardcconsf3d_(const double* xflux, const double* yflux, const double* zflux,
                                  const int* fgi, const int* fgj, const int* fgk,
                                  const double* xdc, const double* ydc, const double* zdc,
                                  const int* dcgi, const int* dcgj, const int* dcgk,
                                  const double* rhs, const int* rgi, const int* rgj,
                                  const int* rgk, const double* sfc, double* soln,
                                  const int* sgi, const int* sgj, const int* sgk,
                                  const int* ifirst, const int* ilast, const int* jfirst,
                                  const int* jlast, const int* kfirst, const int* klast,
                                  const double* dx, const int* offset, const double* maxres) {
  const char* n = "rbgswithfluxmaxvardcconsf3d";
  done(sfac_rbgswithfluxmax3d(xflux, yflux, zflux, *fgi, *fgj, *fgk, xdc, ydc, zdc, *dcgi, *dcgj,
                              *dcgk, 0.0, rhs, *rgi, *rgj, *rgk, nullptr, 0, 0, 0, *sfc, soln,
                              *sgi, *sgj, *sgk, *ifirst, *ilast, *jfirst, *jlast, *kfirst,
                              *klast, dx, *offset, maxres_or_abort(n), nullptr),
       n);
  read_maxres(maxres, n);
}

void rbgswithfluxmaxcondcconsf3d_(const double* xflux, const double* yflux, const double* zflux,
                                  const int* fgi, const int* fgj, const int* fgk,
                                  const double* dc, const double* rhs, const int* rgi,
                                  const int* rgj, const int* rgk, const double* sfc,
                                  double* soln, const int* sgi, const int* sgj, const int* sgk,
                                  const int* ifirst, const int* ilast, const int* jfirst,
                                  const int* jlast, const int* kfirst, const int* klast,
                                  const double* dx, const int* offset, const double* maxres) {
  const char* n = "rbgswithfluxmaxcondcconsf3d";
  done(sfac_rbgswithfluxmax3d(xflux, yflux, zflux, *fgi, *fgj, *fgk, nullptr, nullptr, nullptr,
                              0, 0, 0, *dc, rhs, *rgi, *rgj, *rgk, nullptr, 0, 0, 0, *sfc, soln,
                              *sgi, *sgj, *sgk, *ifirst, *ilast, *jfirst, *jlast, *kfirst,
                              *klast, dx, *offset, maxres_or_abort(n), nullptr),
       n);
  read_maxres(maxres, n);
}

void compresvarsca3d_(const double* xflux, const double* yflux, const double* zflux,
                      const int* fgi, const int* fgj, const int* fgk, const double* rhs,
                      const int* rgi, const int* rgj, const int* rgk, double* residual,
                      const int* residualgi, const int* residualgj, const int* residualgk,
                      const double* sf, const int* sfgi, const int* sfgj, const int* sfgk,
                      const double* soln, const int* sgi, const int* sgj, const int* sgk,
                      const int* ifirst, const int* ilast, const int* jfirst, const int* jlast,
                      const int* kfirst, const int* klast, const double* dx) {
  (void)residualgi;
  (void)residualgj;
  (void)residualgk;
  done(sfac_compres3d(xflux, yflux, zflux, *fgi, *fgj, *fgk, rhs, *rgi, *rgj, *rgk, residual, sf,
                      *sfgi, *sfgj, *sfgk, 0.0, soln, *sgi, *sgj, *sgk, *ifirst, *ilast, *jfirst,
                      *jlast, *kfirst, *klast, dx, nullptr),
       "compresvarsca3d");
}

void compresconsca3d_(const double* xflux, const double* yflux, const double* zflux,
                      const int* fgi, const int* fgj, const int* fgk, const double* rhs,
                      const int* rgi, const int* rgj, const int* rgk, double* residual,
                      const int* residualgi, const int* residualgj, const int* residualgk,
                      const double* sfc, const double* soln, const int* sgi, const int* sgj,
                      const int* sgk, const int* ifirst, const int* ilast, const int* jfirst,
                      const int* jlast, const int* kfirst, const int* klast, const double* dx) {
  (void)residualgi;
  (void)residualgj;
  (void)residualgk;
  done(sfac_compres3d(xflux, yflux, zflux, *fgi, *fgj, *fgk, rhs, *rgi, *rgj, *rgk, residual,
                      nullptr, 0, 0, 0, *sfc, soln, *sgi, *sgj, *sgk, *ifirst, *ilast, *jfirst,
                      *jlast, *kfirst, *klast, dx, nullptr),
       "compresconsca3d");
}

void ewingfixfluxvardc3d_(const double* xflux, const double* yflux, const double* zflux,
                          const int* fgi, const int* fgj, const int* fgk, const double* xdc,
                          const double* ydc, const double* zdc, const int* dcgi, const int* dcgj,
                          const int* dcgk, const double* soln, const int* sgi, const int* sgj,
                          const int* sgk, const int* ifirst, const int* ilast, const int* jfirst,
                          const int* jlast, const int* kfirst, const int* klast,
                          const int* location_index, const int* ratio, const int* blower,
                          const int* bupper, const double* dx) {
  done(sfac_ewingfixflux3d(const_cast<double*>(xflux), const_cast<double*>(yflux),
                           const_cast<double*>(zflux), *fgi, *fgj, *fgk, xdc, ydc, zdc, *dcgi,
                           *dcgj, *dcgk, 0.0, soln, *sgi, *sgj, *sgk, *ifirst, *ilast, *jfirst,
                           *jlast, *kfirst, *klast, *location_index, ratio, blower, bupper, dx,
                           nullptr),
       "ewingfixfluxvardc3d");
}

void ewingfixfluxcondc3d_(const double* xflux, const double* yflux, const double* zflux,
                          const int* fgi, const int* fgj, const int* fgk, const double* dc,
                          const double* soln, const int* sgi, const int* sgj, const int* sgk,
                          const int* ifirst, const int* ilast, const int* jfirst,
                          const int* jlast, const int* kfirst, const int* klast,
                          const int* location_index, const int* ratio, const int* blower,
                          const int* bupper, const double* dx) {
  done(sfac_ewingfixflux3d(const_cast<double*>(xflux), const_cast<double*>(yflux),
                           const_cast<double*>(zflux), *fgi, *fgj, *fgk, nullptr, nullptr,
                           nullptr, 0, 0, 0, *dc, soln, *sgi, *sgj, *sgk, *ifirst, *ilast,
                           *jfirst, *jlast, *kfirst, *klast, *location_index, ratio, blower,
                           bupper, dx, nullptr),
       "ewingfixfluxcondc3d");
}

// ------------------------------------------------ Robin boundary ghosts

void settype1cells2d_(double* data, const int* dil, const int* dih, const int* djl,
                      const int* djh, const double* a, const double* b, const double* g,
                      const int* il, const int* ih, const int* jl, const int* jh,
                      const int* ibeg, const int* iend, const int* jbeg, const int* jend,
                      const int* face, const int* ghos, const int* inte, const int* location,
                      const double* h, const int* zerog) {
  done(sfac_settype1cells2d(data, *dil, *dih, *djl, *djh, a, b, g, *il, *ih, *jl, *jh, *ibeg,
                            *iend, *jbeg, *jend, *face, *ghos, *inte, *location, *h, *zerog,
                            nullptr),
       "settype1cells2d");
}

void settype2cells2d_(double* data, const int* dil, const int* dih, const int* djl,
                      const int* djh, const int* lower, const int* upper, const int* location) {
  done(sfac_settype2cells2d(data, *dil, *dih, *djl, *djh, lower, upper, *location, nullptr),
       "settype2cells2d");
}

void settype1cells3d_(double* data, const int* dil, const int* dih, const int* djl,
                      const int* djh, const int* dkl, const int* dkh, const double* a,
                      const double* b, const double* g, const int* il, const int* ih,
                      const int* jl, const int* jh, const int* kl, const int* kh,
                      const int* ibeg, const int* iend, const int* jbeg, const int* jend,
                      const int* kbeg, const int* kend, const int* face, const int* ghos,
                      const int* inte, const int* location, const double* h, const int* zerog) {
  done(sfac_settype1cells3d(data, *dil, *dih, *djl, *djh, *dkl, *dkh, a, b, g, *il, *ih, *jl,
                            *jh, *kl, *kh, *ibeg, *iend, *jbeg, *jend, *kbeg, *kend, *face,
                            *ghos, *inte, *location, *h, *zerog, nullptr),
       "settype1cells3d");
}

void settype2cells3d_(double* data, const int* dil, const int* dih, const int* djl,
                      const int* djh, const int* dkl, const int* dkh, const int* lower,
                      const int* upper, const int* location) {
  done(sfac_settype2cells3d(data, *dil, *dih, *djl, *djh, *dkl, *dkh, lower, upper, *location,
                            nullptr),
       "settype2cells3d");
}

void settype3cells3d_(double* data, const int* dil, const int* dih, const int* djl,
                      const int* djh, const int* dkl, const int* dkh, const int* lower,
                      const int* upper, const int* location) {
  done(sfac_settype3cells3d(data, *dil, *dih, *djl, *djh, *dkl, *dkh, lower, upper, *location,
                            nullptr),
       "settype3cells3d");
}

// ------------------------------------------------- transfer operators
// (per-axis box corners as separate by-address ints, in the call order of
// geom/CartesianCellDoubleWeightedAverage.C:183-246,
// geom/CartesianCellDoubleConservativeLinearRefine.C:263-378,
// geom/CartesianCellDoubleLinearRefine.C:212-293 and
// pdat/CellDoubleConstantRefine.C:154-176; the conservative refine's
// difference and slope work arrays are not needed and are left untouched)

void cartwgtavgcelldoub2d_(const int* ifc0, const int* ifc1, const int* ilc0, const int* ilc1,
                           const int* filo0, const int* filo1, const int* fihi0,
                           const int* fihi1, const int* cilo0, const int* cilo1,
                           const int* cihi0, const int* cihi1, const int* ratio,
                           const double* dxf, const double* dxc, const double* arrayf,
                           double* arrayc) {
  const int ifc[2] = {*ifc0, *ifc1}, ilc[2] = {*ilc0, *ilc1}, filo[2] = {*filo0, *filo1},
            fihi[2] = {*fihi0, *fihi1}, cilo[2] = {*cilo0, *cilo1}, cihi[2] = {*cihi0, *cihi1};
  done(sfac_cartwgtavgcelldoub(2, ifc, ilc, filo, fihi, cilo, cihi, ratio, dxf, dxc, arrayf,
                               arrayc, nullptr),
       "cartwgtavgcelldoub2d");
}

void cartwgtavgcelldoub3d_(const int* ifc0, const int* ifc1, const int* ifc2, const int* ilc0,
                           const int* ilc1, const int* ilc2, const int* filo0,
                           const int* filo1, const int* filo2, const int* fihi0,
                           const int* fihi1, const int* fihi2, const int* cilo0,
                           const int* cilo1, const int* cilo2, const int* cihi0,
                           const int* cihi1, const int* cihi2, const int* ratio,
                           const double* dxf, const double* dxc, const double* arrayf,
                           double* arrayc) {
  const int ifc[3] = {*ifc0, *ifc1, *ifc2}, ilc[3] = {*ilc0, *ilc1, *ilc2},
            filo[3] = {*filo0, *filo1, *filo2}, fihi[3] = {*fihi0, *fihi1, *fihi2},
            cilo[3] = {*cilo0, *cilo1, *cilo2}, cihi[3] = {*cihi0, *cihi1, *cihi2};
  done(sfac_cartwgtavgcelldoub(3, ifc, ilc, filo, fihi, cilo, cihi, ratio, dxf, dxc, arrayf,
                               arrayc, nullptr),
       "cartwgtavgcelldoub3d");
}

void cartclinrefcelldoub2d_(const int* ifc0, const int* ifc1, const int* ilc0, const int* ilc1,
                            const int* iff0, const int* iff1, const int* ilf0, const int* ilf1,
                            const int* cilo0, const int* cilo1, const int* cihi0,
                            const int* cihi1, const int* filo0, const int* filo1,
                            const int* fihi0, const int* fihi1, const int* ratio,
                            const double* dxc, const double* dxf, const double* arrayc,
                            double* arrayf, double* diff0, double* slope0, double* diff1,
                            double* slope1) {
  (void)diff0;
  (void)slope0;
  (void)diff1;
  (void)slope1;
  const int ifc[2] = {*ifc0, *ifc1}, ilc[2] = {*ilc0, *ilc1}, iff[2] = {*iff0, *iff1},
            ilf[2] = {*ilf0, *ilf1}, cilo[2] = {*cilo0, *cilo1}, cihi[2] = {*cihi0, *cihi1},
            filo[2] = {*filo0, *filo1}, fihi[2] = {*fihi0, *fihi1};
  done(sfac_cartclinrefcelldoub(2, ifc, ilc, iff, ilf, cilo, cihi, filo, fihi, ratio, dxc, dxf,
                                arrayc, arrayf, nullptr),
       "cartclinrefcelldoub2d");
}

void cartclinrefcelldoub3d_(const int* ifc0, const int* ifc1, const int* ifc2, const int* ilc0,
                            const int* ilc1, const int* ilc2, const int* iff0, const int* iff1,
                            const int* iff2, const int* ilf0, const int* ilf1, const int* ilf2,
                            const int* cilo0, const int* cilo1, const int* cilo2,
                            const int* cihi0, const int* cihi1, const int* cihi2,
                            const int* filo0, const int* filo1, const int* filo2,
                            const int* fihi0, const int* fihi1, const int* fihi2,
                            const int* ratio, const double* dxc, const double* dxf,
                            const double* arrayc, double* arrayf, double* diff0,
                            double* slope0, double* diff1, double* slope1, double* diff2,
                            double* slope2) {
  (void)diff0;
  (void)slope0;
  (void)diff1;
  (void)slope1;
  (void)diff2;
  (void)slope2;
  const int ifc[3] = {*ifc0, *ifc1, *ifc2}, ilc[3] = {*ilc0, *ilc1, *ilc2},
            iff[3] = {*iff0, *iff1, *iff2}, ilf[3] = {*ilf0, *ilf1, *ilf2},
            cilo[3] = {*cilo0, *cilo1, *cilo2}, cihi[3] = {*cihi0, *cihi1, *cihi2},
            filo[3] = {*filo0, *filo1, *filo2}, fihi[3] = {*fihi0, *fihi1, *fihi2};
  done(sfac_cartclinrefcelldoub(3, ifc, ilc, iff, ilf, cilo, cihi, filo, fihi, ratio, dxc, dxf,
                                arrayc, arrayf, nullptr),
       "cartclinrefcelldoub3d");
}

void cartlinrefcelldoub2d_(const int* ifc0, const int* ifc1, const int* ilc0, const int* ilc1,
                           const int* iff0, const int* iff1, const int* ilf0, const int* ilf1,
                           const int* cilo0, const int* cilo1, const int* cihi0,
                           const int* cihi1, const int* filo0, const int* filo1,
                           const int* fihi0, const int* fihi1, const int* ratio,
                           const double* dxc, const double* dxf, const double* arrayc,
                           double* arrayf) {
  const int ifc[2] = {*ifc0, *ifc1}, ilc[2] = {*ilc0, *ilc1}, iff[2] = {*iff0, *iff1},
            ilf[2] = {*ilf0, *ilf1}, cilo[2] = {*cilo0, *cilo1}, cihi[2] = {*cihi0, *cihi1},
            filo[2] = {*filo0, *filo1}, fihi[2] = {*fihi0, *fihi1};
  done(sfac_cartlinrefcelldoub(2, ifc, ilc, iff, ilf, cilo, cihi, filo, fihi, ratio, dxc, dxf,
                               arrayc, arrayf, nullptr),
       "cartlinrefcelldoub2d");
}

void cartlinrefcelldoub3d_(const int* ifc0, const int* ifc1, const int* ifc2, const int* ilc0,
                           const int* ilc1, const int* ilc2, const int* iff0, const int* iff1,
                           const int* iff2, const int* ilf0, const int* ilf1, const int* ilf2,
                           const int* cilo0, const int* cilo1, const int* cilo2,
                           const int* cihi0, const int* cihi1, const int* cihi2,
                           const int* filo0, const int* filo1, const int* filo2,
                           const int* fihi0, const int* fihi1, const int* fihi2,
                           const int* ratio, const double* dxc, const double* dxf,
                           const double* arrayc, double* arrayf) {
  const int ifc[3] = {*ifc0, *ifc1, *ifc2}, ilc[3] = {*ilc0, *ilc1, *ilc2},
            iff[3] = {*iff0, *iff1, *iff2}, ilf[3] = {*ilf0, *ilf1, *ilf2},
            cilo[3] = {*cilo0, *cilo1, *cilo2}, cihi[3] = {*cihi0, *cihi1, *cihi2},
            filo[3] = {*filo0, *filo1, *filo2}, fihi[3] = {*fihi0, *fihi1, *fihi2};
  done(sfac_cartlinrefcelldoub(3, ifc, ilc, iff, ilf, cilo, cihi, filo, fihi, ratio, dxc, dxf,
                               arrayc, arrayf, nullptr),
       "cartlinrefcelldoub3d");
}

void conrefcelldoub2d_(const int* ifc0, const int* ifc1, const int* ilc0, const int* ilc1,
                       const int* iff0, const int* iff1, const int* ilf0, const int* ilf1,
                       const int* cilo0, const int* cilo1, const int* cihi0, const int* cihi1,
                       const int* filo0, const int* filo1, const int* fihi0, const int* fihi1,
                       const int* ratio, const double* arrayc, double* arrayf) {
  (void)ifc0;
  (void)ifc1;
  (void)ilc0;
  (void)ilc1;
  const int iff[2] = {*iff0, *iff1}, ilf[2] = {*ilf0, *ilf1}, cilo[2] = {*cilo0, *cilo1},
            cihi[2] = {*cihi0, *cihi1}, filo[2] = {*filo0, *filo1}, fihi[2] = {*fihi0, *fihi1};
  done(sfac_conrefcelldoub(2, iff, ilf, cilo, cihi, filo, fihi, ratio, arrayc, arrayf, nullptr),
       "conrefcelldoub2d");
}

void conrefcelldoub3d_(const int* ifc0, const int* ifc1, const int* ifc2, const int* ilc0,
                       const int* ilc1, const int* ilc2, const int* iff0, const int* iff1,
                       const int* iff2, const int* ilf0, const int* ilf1, const int* ilf2,
                       const int* cilo0, const int* cilo1, const int* cilo2, const int* cihi0,
                       const int* cihi1, const int* cihi2, const int* filo0, const int* filo1,
                       const int* filo2, const int* fihi0, const int* fihi1, const int* fihi2,
                       const int* ratio, const double* arrayc, double* arrayf) {
  (void)ifc0;
  (void)ifc1;
  (void)ifc2;
  (void)ilc0;
  (void)ilc1;
  (void)ilc2;
  const int iff[3] = {*iff0, *iff1, *iff2}, ilf[3] = {*ilf0, *ilf1, *ilf2},
            cilo[3] = {*cilo0, *cilo1, *cilo2}, cihi[3] = {*cihi0, *cihi1, *cihi2},
            filo[3] = {*filo0, *filo1, *filo2}, fihi[3] = {*fihi0, *fihi1, *fihi2};
  done(sfac_conrefcelldoub(3, iff, ilf, cilo, cihi, filo, fihi, ratio, arrayc, arrayf, nullptr),
       "conrefcelldoub3d");
}

// outerside flux average onto the coarse faces of axis 0, 1 (2-D) or 0, 1,
// 2 (3-D), in the call order of
// geom/CartesianOutersideDoubleWeightedAverage.C:242-400
void cartwgtavgoutsidedoub2d0_(const int* ifc0, const int* ifc1, const int* ilc0,
                               const int* ilc1, const int* filo0, const int* filo1,
                               const int* fihi0, const int* fihi1, const int* cilo0,
                               const int* cilo1, const int* cihi0, const int* cihi1,
                               const int* ratio, const double* dxf, const double* dxc,
                               const double* arrayf, double* arrayc);
void cartwgtavgoutsidedoub2d1_(const int* ifc0, const int* ifc1, const int* ilc0,
                               const int* ilc1, const int* filo0, const int* filo1,
                               const int* fihi0, const int* fihi1, const int* cilo0,
                               const int* cilo1, const int* cihi0, const int* cihi1,
                               const int* ratio, const double* dxf, const double* dxc,
                               const double* arrayf, double* arrayc);
void cartwgtavgoutsidedoub3d0_(const int* ifc0, const int* ifc1, const int* ifc2,
                               const int* ilc0, const int* ilc1, const int* ilc2,
                               const int* filo0, const int* filo1, const int* filo2,
                               const int* fihi0, const int* fihi1, const int* fihi2,
                               const int* cilo0, const int* cilo1, const int* cilo2,
                               const int* cihi0, const int* cihi1, const int* cihi2,
                               const int* ratio, const double* dxf, const double* dxc,
                               const double* arrayf, double* arrayc);
void cartwgtavgoutsidedoub3d1_(const int* ifc0, const int* ifc1, const int* ifc2,
                               const int* ilc0, const int* ilc1, const int* ilc2,
                               const int* filo0, const int* filo1, const int* filo2,
                               const int* fihi0, const int* fihi1, const int* fihi2,
                               const int* cilo0, const int* cilo1, const int* cilo2,
                               const int* cihi0, const int* cihi1, const int* cihi2,
                               const int* ratio, const double* dxf, const double* dxc,
                               const double* arrayf, double* arrayc);
void cartwgtavgoutsidedoub3d2_(const int* ifc0, const int* ifc1, const int* ifc2,
                               const int* ilc0, const int* ilc1, const int* ilc2,
                               const int* filo0, const int* filo1, const int* filo2,
                               const int* fihi0, const int* fihi1, const int* fihi2,
                               const int* cilo0, const int* cilo1, const int* cilo2,
                               const int* cihi0, const int* cihi1, const int* cihi2,
                               const int* ratio, const double* dxf, const double* dxc,
                               const double* arrayf, double* arrayc);

}  // extern "C"

namespace {

void outside2(int axis, const int* ifc0, const int* ifc1, const int* ilc0, const int* ilc1,
              const int* filo0, const int* filo1, const int* fihi0, const int* fihi1,
              const int* cilo0, const int* cilo1, const int* cihi0, const int* cihi1,
              const int* ratio, const double* dxf, const double* dxc, const double* arrayf,
              double* arrayc, const char* name) {
  const int ifc[2] = {*ifc0, *ifc1}, ilc[2] = {*ilc0, *ilc1}, filo[2] = {*filo0, *filo1},
            fihi[2] = {*fihi0, *fihi1}, cilo[2] = {*cilo0, *cilo1}, cihi[2] = {*cihi0, *cihi1};
  done(sfac_cartwgtavgoutsidedoub(2, axis, ifc, ilc, filo, fihi, cilo, cihi, ratio, dxf, dxc,
                                  arrayf, arrayc, nullptr),
       name);
}

void outside3(int axis, const int* ifc0, const int* ifc1, const int* ifc2, const int* ilc0,
              const int* ilc1, const int* ilc2, const int* filo0, const int* filo1,
              const int* filo2, const int* fihi0, const int* fihi1, const int* fihi2,
              const int* cilo0, const int* cilo1, const int* cilo2, const int* cihi0,
              const int* cihi1, const int* cihi2, const int* ratio, const double* dxf,
              const double* dxc, const double* arrayf, double* arrayc, const char* name) {
  const int ifc[3] = {*ifc0, *ifc1, *ifc2}, ilc[3] = {*ilc0, *ilc1, *ilc2},
            filo[3] = {*filo0, *filo1, *filo2}, fihi[3] = {*fihi0, *fihi1, *fihi2},
            cilo[3] = {*cilo0, *cilo1, *cilo2}, cihi[3] = {*cihi0, *cihi1, *cihi2};
  done(sfac_cartwgtavgoutsidedoub(3, axis, ifc, ilc, filo, fihi, cilo, cihi, ratio, dxf, dxc,
                                  arrayf, arrayc, nullptr),
       name);
}

}  // namespace

#define SFAC_OUTSIDE2(AX)                                                                     \
  extern "C" void cartwgtavgoutsidedoub2d##AX##_(                                             \
      const int* a0, const int* a1, const int* b0, const int* b1, const int* c0,              \
      const int* c1, const int* d0, const int* d1, const int* e0, const int* e1,              \
      const int* f0, const int* f1, const int* ratio, const double* dxf, const double* dxc,   \
      const double* arrayf, double* arrayc) {                                                 \
    outside2(AX, a0, a1, b0, b1, c0, c1, d0, d1, e0, e1, f0, f1, ratio, dxf, dxc, arrayf,     \
             arrayc, "cartwgtavgoutsidedoub2d" #AX);                                          \
  }
#define SFAC_OUTSIDE3(AX)                                                                     \
  extern "C" void cartwgtavgoutsidedoub3d##AX##_(                                             \
      const int* a0, const int* a1, const int* a2, const int* b0, const int* b1,              \
      const int* b2, const int* c0, const int* c1, const int* c2, const int* d0,              \
      const int* d1, const int* d2, const int* e0, const int* e1, const int* e2,              \
      const int* f0, const int* f1, const int* f2, const int* ratio, const double* dxf,       \
      const double* dxc, const double* arrayf, double* arrayc) {                              \
    outside3(AX, a0, a1, a2, b0, b1, b2, c0, c1, c2, d0, d1, d2, e0, e1, e2, f0, f1, f2,      \
             ratio, dxf, dxc, arrayf, arrayc, "cartwgtavgoutsidedoub3d" #AX);                 \
  }
SFAC_OUTSIDE2(0)
SFAC_OUTSIDE2(1)
SFAC_OUTSIDE3(0)
SFAC_OUTSIDE3(1)
SFAC_OUTSIDE3(2)
