/* samrai_f77.h -- link-compatible twins of the Fortran kernels that SAMRAI's
 * FAC operator and Robin boundary helper call through SAMRAI_F77_FUNC
 * (lower case, trailing underscore, every argument by address):
 * solv/CellPoissonFACOps.C:62-543, solv/CartesianRobinBcHelper.C:32-73 and
 * the transfer operators' calls (geom/CartesianCellDoubleWeightedAverage.C,
 * geom/CartesianOutersideDoubleWeightedAverage.C,
 * geom/CartesianCellDoubleConservativeLinearRefine.C,
 * geom/CartesianCellDoubleLinearRefine.C, pdat/CellDoubleConstantRefine.C;
 * box corners one by-address int per axis).  The C++ declarations there take
 * some scalars as `const double&` / `const int&`, which the C ABI passes as
 * the pointers declared here.
 *
 * Arrays are device pointers into ArrayData storage (the reference
 * allocates it through tbox::ResourceAllocator / Umpire,
 * pdat/ArrayData.C:81-112); scalars, index vectors (dx, ratio, box corners)
 * and the red-black routines' maxres are host memory.  Each routine runs its
 * kernel on the null stream and returns when it has finished, as the Fortran
 * does.  There is no status: an error prints the library's last error and
 * aborts, as TBOX_ERROR does.  The per-patch sfac_ entry points of
 * samrai_fac.h are the stream-ordered, status-returning form of the same
 * kernels.
 */
#ifndef SAMRAI_F77_H
#define SAMRAI_F77_H

#ifdef __cplusplus
extern "C" {
#endif

void compfluxvardc2d_(double* xflux, double* yflux, const int* fgi, const int* fgj,
    const double* xdc, const double* ydc, const int* dcgi, const int* dcgj,
    const double* soln, const int* sgi, const int* sgj, const int* ifirst, const int* ilast,
    const int* jfirst, const int* jlast, const double* dx);
void compfluxcondc2d_(double* xflux, double* yflux, const int* fgi, const int* fgj,
    const double* dc, const double* soln, const int* sgi, const int* sgj, const int* ifirst,
    const int* ilast, const int* jfirst, const int* jlast, const double* dx);
void rbgswithfluxmaxvardcvarsf2d_(const double* xflux, const double* yflux, const int* fgi,
    const int* fgj, const double* xdc, const double* ydc, const int* dcgi, const int* dcgj,
    const double* rhs, const int* rgi, const int* rgj, const double* sf, const int* sfgi,
    const int* sfgj, double* soln, const int* sgi, const int* sgj, const int* ifirst,
    const int* ilast, const int* jfirst, const int* jlast, const double* dx,
    const int* offset, const double* maxres);
void rbgswithfluxmaxcondcvarsf2d_(const double* xflux, const double* yflux, const int* fgi,
    const int* fgj, const double* dc, const double* rhs, const int* rgi, const int* rgj,
    const double* sf, const int* sfgi, const int* sfgj, double* soln, const int* sgi,
    const int* sgj, const int* ifirst, const int* ilast, const int* jfirst, const int* jlast,
    const double* dx, const int* offset, const double* maxres);
void rbgswithfluxmaxvardcconsf2d_(const double* xflux, const double* yflux, const int* fgi,
    const int* fgj, const double* xdc, const double* ydc, const int* dcgi, const int* dcgj,
    const double* rhs, const int* rgi, const int* rgj, const double* sfc, double* soln,
    const int* sgi, const int* sgj, const int* ifirst, const int* ilast, const int* jfirst,
    const int* jlast, const double* dx, const int* offset, const double* maxres);
void rbgswithfluxmaxcondcconsf2d_(const double* xflux, const double* yflux, const int* fgi,
    const int* fgj, const double* dc, const double* rhs, const int* rgi, const int* rgj,
    const double* sfc, double* soln, const int* sgi, const int* sgj, const int* ifirst,
    const int* ilast, const int* jfirst, const int* jlast, const double* dx,
    const int* offset, const double* maxres);
void compresvarsca2d_(const double* xflux, const double* yflux, const int* fgi,
    const int* fgj, const double* rhs, const int* rgi, const int* rgj, double* residual,
    const int* residualgi, const int* residualgj, const double* sf, const int* sfgi,
    const int* sfgj, const double* soln, const int* sgi, const int* sgj, const int* ifirst,
    const int* ilast, const int* jfirst, const int* jlast, const double* dx);
void compresconsca2d_(const double* xflux, const double* yflux, const int* fgi,
    const int* fgj, const double* rhs, const int* rgi, const int* rgj, double* residual,
    const int* residualgi, const int* residualgj, const double* sfc, const double* soln,
    const int* sgi, const int* sgj, const int* ifirst, const int* ilast, const int* jfirst,
    const int* jlast, const double* dx);
void ewingfixfluxvardc2d_(const double* xflux, const double* yflux, const int* fgi,
    const int* fgj, const double* xdc, const double* ydc, const int* dcgi, const int* dcgj,
    const double* soln, const int* sgi, const int* sgj, const int* ifirst, const int* ilast,
    const int* jfirst, const int* jlast, const int* location_index, const int* ratio,
    const int* blower, const int* bupper, const double* dx);
void ewingfixfluxcondc2d_(const double* xflux, const double* yflux, const int* fgi,
    const int* fgj, const double* dc, const double* soln, const int* sgi, const int* sgj,
    const int* ifirst, const int* ilast, const int* jfirst, const int* jlast,
    const int* location_index, const int* ratio, const int* blower, const int* bupper,
    const double* dx);
void compfluxvardc3d_(double* xflux, double* yflux, double* zflux, const int* fgi,
    const int* fgj, const int* fgk, const double* xdc, const double* ydc, const double* zdc,
    const int* dcgi, const int* dcgj, const int* dcgk, const double* soln, const int* sgi,
    const int* sgj, const int* sgk, const int* ifirst, const int* ilast, const int* jfirst,
    const int* jlast, const int* kfirst, const int* klast, const double* dx);
void compfluxcondc3d_(double* xflux, double* yflux, double* zflux, const int* fgi,
    const int* fgj, const int* fgk, const double* dc, const double* soln, const int* sgi,
    const int* sgj, const int* sgk, const int* ifirst, const int* ilast, const int* jfirst,
    const int* jlast, const int* kfirst, const int* klast, const double* dx);
void rbgswithfluxmaxvardcvarsf3d_(const double* xflux, const double* yflux,
    const double* zflux, const int* fgi, const int* fgj, const int* fgk, const double* xdc,
    const double* ydc, const double* zdc, const int* dcgi, const int* dcgj, const int* dcgk,
    const double* rhs, const int* rgi, const int* rgj, const int* rgk, const double* sf,
    const int* sfgi, const int* sfgj, const int* sfgk, double* soln, const int* sgi,
    const int* sgj, const int* sgk, const int* ifirst, const int* ilast, const int* jfirst,
    const int* jlast, const int* kfirst, const int* klast, const double* dx,
    const int* offset, const double* maxres);
void rbgswithfluxmaxcondcvarsf3d_(const double* xflux, const double* yflux,
    const double* zflux, const int* fgi, const int* fgj, const int* fgk, const double* dc,
    const double* rhs, const int* rgi, const int* rgj, const int* rgk, const double* sf,
    const int* sfgi, const int* sfgj, const int* sfgk, double* soln, const int* sgi,
    const int* sgj, const int* sgk, const int* ifirst, const int* ilast, const int* jfirst,
    const int* jlast, const int* kfirst, const int* klast, const double* dx,
    const int* offset, const double* maxres);
void rbgswithfluxmaxvardcconsf3d_(const double* xflux, const double* yflux,
    const double* zflux, const int* fgi, const int* fgj, const int* fgk, const double* xdc,
    const double* ydc, const double* zdc, const int* dcgi, const int* dcgj, const int* dcgk,
    const double* rhs, const int* rgi, const int* rgj, const int* rgk, const double* sfc,
    double* soln, const int* sgi, const int* sgj, const int* sgk, const int* ifirst,
    const int* ilast, const int* jfirst, const int* jlast, const int* kfirst,
    const int* klast, const double* dx, const int* offset, const double* maxres);
void rbgswithfluxmaxcondcconsf3d_(const double* xflux, const double* yflux,
    const double* zflux, const int* fgi, const int* fgj, const int* fgk, const double* dc,
    const double* rhs, const int* rgi, const int* rgj, const int* rgk, const double* sfc,
    double* soln, const int* sgi, const int* sgj, const int* sgk, const int* ifirst,
    const int* ilast, const int* jfirst, const int* jlast, const int* kfirst,
    const int* klast, const double* dx, const int* offset, const double* maxres);
void compresvarsca3d_(const double* xflux, const double* yflux, const double* zflux,
    const int* fgi, const int* fgj, const int* fgk, const double* rhs, const int* rgi,
    const int* rgj, const int* rgk, double* residual, const int* residualgi,
    const int* residualgj, const int* residualgk, const double* sf, const int* sfgi,
    const int* sfgj, const int* sfgk, const double* soln, const int* sgi, const int* sgj,
    const int* sgk, const int* ifirst, const int* ilast, const int* jfirst, const int* jlast,
    const int* kfirst, const int* klast, const double* dx);
void compresconsca3d_(const double* xflux, const double* yflux, const double* zflux,
    const int* fgi, const int* fgj, const int* fgk, const double* rhs, const int* rgi,
    const int* rgj, const int* rgk, double* residual, const int* residualgi,
    const int* residualgj, const int* residualgk, const double* sfc, const double* soln,
    const int* sgi, const int* sgj, const int* sgk, const int* ifirst, const int* ilast,
    const int* jfirst, const int* jlast, const int* kfirst, const int* klast, const double* dx);
void ewingfixfluxvardc3d_(const double* xflux, const double* yflux, const double* zflux,
    const int* fgi, const int* fgj, const int* fgk, const double* xdc, const double* ydc,
    const double* zdc, const int* dcgi, const int* dcgj, const int* dcgk, const double* soln,
    const int* sgi, const int* sgj, const int* sgk, const int* ifirst, const int* ilast,
    const int* jfirst, const int* jlast, const int* kfirst, const int* klast,
    const int* location_index, const int* ratio, const int* blower, const int* bupper,
    const double* dx);
void ewingfixfluxcondc3d_(const double* xflux, const double* yflux, const double* zflux,
    const int* fgi, const int* fgj, const int* fgk, const double* dc, const double* soln,
    const int* sgi, const int* sgj, const int* sgk, const int* ifirst, const int* ilast,
    const int* jfirst, const int* jlast, const int* kfirst, const int* klast,
    const int* location_index, const int* ratio, const int* blower, const int* bupper,
    const double* dx);
void settype1cells2d_(double* data, const int* dil, const int* dih, const int* djl,
    const int* djh, const double* a, const double* b, const double* g, const int* il,
    const int* ih, const int* jl, const int* jh, const int* ibeg, const int* iend,
    const int* jbeg, const int* jend, const int* face, const int* ghos, const int* inte,
    const int* location, const double* h, const int* zerog);
void settype2cells2d_(double* data, const int* dil, const int* dih, const int* djl,
    const int* djh, const int* lower, const int* upper, const int* location);
void settype1cells3d_(double* data, const int* dil, const int* dih, const int* djl,
    const int* djh, const int* dkl, const int* dkh, const double* a, const double* b,
    const double* g, const int* il, const int* ih, const int* jl, const int* jh,
    const int* kl, const int* kh, const int* ibeg, const int* iend, const int* jbeg,
    const int* jend, const int* kbeg, const int* kend, const int* face, const int* ghos,
    const int* inte, const int* location, const double* h, const int* zerog);
void settype2cells3d_(double* data, const int* dil, const int* dih, const int* djl,
    const int* djh, const int* dkl, const int* dkh, const int* lower, const int* upper,
    const int* location);
void settype3cells3d_(double* data, const int* dil, const int* dih, const int* djl,
    const int* djh, const int* dkl, const int* dkh, const int* lower, const int* upper,
    const int* location);
void cartwgtavgcelldoub2d_(const int* ifc0, const int* ifc1, const int* ilc0, const int* ilc1,
    const int* filo0, const int* filo1, const int* fihi0, const int* fihi1, const int* cilo0,
    const int* cilo1, const int* cihi0, const int* cihi1, const int* ratio, const double* dxf,
    const double* dxc, const double* arrayf, double* arrayc);
void cartwgtavgcelldoub3d_(const int* ifc0, const int* ifc1, const int* ifc2, const int* ilc0,
    const int* ilc1, const int* ilc2, const int* filo0, const int* filo1, const int* filo2,
    const int* fihi0, const int* fihi1, const int* fihi2, const int* cilo0, const int* cilo1,
    const int* cilo2, const int* cihi0, const int* cihi1, const int* cihi2, const int* ratio,
    const double* dxf, const double* dxc, const double* arrayf, double* arrayc);
void cartclinrefcelldoub2d_(const int* ifc0, const int* ifc1, const int* ilc0,
    const int* ilc1, const int* iff0, const int* iff1, const int* ilf0, const int* ilf1,
    const int* cilo0, const int* cilo1, const int* cihi0, const int* cihi1, const int* filo0,
    const int* filo1, const int* fihi0, const int* fihi1, const int* ratio, const double* dxc,
    const double* dxf, const double* arrayc, double* arrayf, double* diff0, double* slope0,
    double* diff1, double* slope1);
void cartclinrefcelldoub3d_(const int* ifc0, const int* ifc1, const int* ifc2,
    const int* ilc0, const int* ilc1, const int* ilc2, const int* iff0, const int* iff1,
    const int* iff2, const int* ilf0, const int* ilf1, const int* ilf2, const int* cilo0,
    const int* cilo1, const int* cilo2, const int* cihi0, const int* cihi1, const int* cihi2,
    const int* filo0, const int* filo1, const int* filo2, const int* fihi0, const int* fihi1,
    const int* fihi2, const int* ratio, const double* dxc, const double* dxf,
    const double* arrayc, double* arrayf, double* diff0, double* slope0, double* diff1,
    double* slope1, double* diff2, double* slope2);
void cartlinrefcelldoub2d_(const int* ifc0, const int* ifc1, const int* ilc0, const int* ilc1,
    const int* iff0, const int* iff1, const int* ilf0, const int* ilf1, const int* cilo0,
    const int* cilo1, const int* cihi0, const int* cihi1, const int* filo0, const int* filo1,
    const int* fihi0, const int* fihi1, const int* ratio, const double* dxc,
    const double* dxf, const double* arrayc, double* arrayf);
void cartlinrefcelldoub3d_(const int* ifc0, const int* ifc1, const int* ifc2, const int* ilc0,
    const int* ilc1, const int* ilc2, const int* iff0, const int* iff1, const int* iff2,
    const int* ilf0, const int* ilf1, const int* ilf2, const int* cilo0, const int* cilo1,
    const int* cilo2, const int* cihi0, const int* cihi1, const int* cihi2, const int* filo0,
    const int* filo1, const int* filo2, const int* fihi0, const int* fihi1, const int* fihi2,
    const int* ratio, const double* dxc, const double* dxf, const double* arrayc,
    double* arrayf);
void conrefcelldoub2d_(const int* ifc0, const int* ifc1, const int* ilc0, const int* ilc1,
    const int* iff0, const int* iff1, const int* ilf0, const int* ilf1, const int* cilo0,
    const int* cilo1, const int* cihi0, const int* cihi1, const int* filo0, const int* filo1,
    const int* fihi0, const int* fihi1, const int* ratio, const double* arrayc, double* arrayf);
void conrefcelldoub3d_(const int* ifc0, const int* ifc1, const int* ifc2, const int* ilc0,
    const int* ilc1, const int* ilc2, const int* iff0, const int* iff1, const int* iff2,
    const int* ilf0, const int* ilf1, const int* ilf2, const int* cilo0, const int* cilo1,
    const int* cilo2, const int* cihi0, const int* cihi1, const int* cihi2, const int* filo0,
    const int* filo1, const int* filo2, const int* fihi0, const int* fihi1, const int* fihi2,
    const int* ratio, const double* arrayc, double* arrayf);
void cartwgtavgoutsidedoub2d0_(const int* ifc0, const int* ifc1, const int* ilc0,
    const int* ilc1, const int* filo0, const int* filo1, const int* fihi0, const int* fihi1,
    const int* cilo0, const int* cilo1, const int* cihi0, const int* cihi1, const int* ratio,
    const double* dxf, const double* dxc, const double* arrayf, double* arrayc);
void cartwgtavgoutsidedoub2d1_(const int* ifc0, const int* ifc1, const int* ilc0,
    const int* ilc1, const int* filo0, const int* filo1, const int* fihi0, const int* fihi1,
    const int* cilo0, const int* cilo1, const int* cihi0, const int* cihi1, const int* ratio,
    const double* dxf, const double* dxc, const double* arrayf, double* arrayc);
void cartwgtavgoutsidedoub3d0_(const int* ifc0, const int* ifc1, const int* ifc2,
    const int* ilc0, const int* ilc1, const int* ilc2, const int* filo0, const int* filo1,
    const int* filo2, const int* fihi0, const int* fihi1, const int* fihi2, const int* cilo0,
    const int* cilo1, const int* cilo2, const int* cihi0, const int* cihi1, const int* cihi2,
    const int* ratio, const double* dxf, const double* dxc, const double* arrayf,
    double* arrayc);
void cartwgtavgoutsidedoub3d1_(const int* ifc0, const int* ifc1, const int* ifc2,
    const int* ilc0, const int* ilc1, const int* ilc2, const int* filo0, const int* filo1,
    const int* filo2, const int* fihi0, const int* fihi1, const int* fihi2, const int* cilo0,
    const int* cilo1, const int* cilo2, const int* cihi0, const int* cihi1, const int* cihi2,
    const int* ratio, const double* dxf, const double* dxc, const double* arrayf,
    double* arrayc);
void cartwgtavgoutsidedoub3d2_(const int* ifc0, const int* ifc1, const int* ifc2,
    const int* ilc0, const int* ilc1, const int* ilc2, const int* filo0, const int* filo1,
    const int* filo2, const int* fihi0, const int* fihi1, const int* fihi2, const int* cilo0,
    const int* cilo1, const int* cilo2, const int* cihi0, const int* cihi1, const int* cihi2,
    const int* ratio, const double* dxf, const double* dxc, const double* arrayf,
    double* arrayc);

#ifdef __cplusplus
}
#endif

#endif /* SAMRAI_F77_H */
