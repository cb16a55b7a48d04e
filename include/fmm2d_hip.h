/*
 * fmm2d_hip.h -- C ABI of the MI355X 2D FMM engine (libfmm2d_hip.so).
 *
 * Drop-in entry points: the gfortran-ABI symbols of the reference C
 * translation (lowercase, trailing underscore, every argument by pointer,
 * column-major arrays with nd fastest; c_translation/include/fmm2d_c.h:35-61).
 * Each declaration cites the reference routine it replaces.  Argument
 * meaning and output layout are unchanged:
 *   - outputs are overwritten, unrequested outputs are not touched;
 *   - *iper is written 0; timeinfo[8] receives per-step device times (s);
 *   - *ier: 0 ok, 4 device allocation failure, 13 Helmholtz expansion order
 *     above 10000 (reference meanings), and this engine's own codes
 *       101 HIP runtime error          102 unsupported argument
 *       104 no usable device
 *       106 expansion order beyond the kernels' limits (Laplace 63,
 *           Helmholtz 2000, and the order of a Helmholtz leaf whose
 *           formation tile would exceed 64 KB of LDS).  (103 and 105 are
 *           no longer returned: refinement stops at the keys' 29 levels, and
 *           boxes above 16 wavelengths take the diagonal form.)
 *     The library never aborts or prints.
 * Calls are synchronous and thread-safe (one HIP stream per calling thread).
 */
#ifndef FMM2D_HIP_H
#define FMM2D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> fmm2d_cplx;
extern "C" {
#else
typedef double _Complex fmm2d_cplx;
#endif

typedef int32_t fmm2d_int;

/* ---- Laplace, Cauchy form: pot = sum c log|z| + d/z, grad = d/dz, hess = d2/dz2 */

/* replaces cfmm2d_ (c_translation/src/cfmm2d.c:1053; leaf size from lndiv2d,
 * src/laplace/lndiv2d.f:17-35) */
void cfmm2d_(const fmm2d_int *nd, const double *eps, const fmm2d_int *ns,
             const double *sources, const fmm2d_int *ifcharge, const fmm2d_cplx *charge,
             const fmm2d_int *ifdipole, const fmm2d_cplx *dipstr, fmm2d_int *iper,
             const fmm2d_int *ifpgh, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess,
             const fmm2d_int *nt, const double *targ, const fmm2d_int *ifpghtarg,
             fmm2d_cplx *pottarg, fmm2d_cplx *gradtarg, fmm2d_cplx *hesstarg, fmm2d_int *ier);

/* replaces cfmm2d_ndiv_ (c_translation/src/cfmm2d_ndiv.c:94) */
void cfmm2d_ndiv_(const fmm2d_int *nd, const double *eps, const fmm2d_int *ns,
                  const double *sources, const fmm2d_int *ifcharge, const fmm2d_cplx *charge,
                  const fmm2d_int *ifdipole, const fmm2d_cplx *dipstr, fmm2d_int *iper,
                  const fmm2d_int *ifpgh, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess,
                  const fmm2d_int *nt, const double *targ, const fmm2d_int *ifpghtarg,
                  fmm2d_cplx *pottarg, fmm2d_cplx *gradtarg, fmm2d_cplx *hesstarg,
                  const fmm2d_int *ndiv, const fmm2d_int *idivflag, const fmm2d_int *ifnear,
                  double *timeinfo, fmm2d_int *ier);

/* ---- Laplace, log kernel with complex densities: grad (nd,2,n), hess (nd,3,n) */

/* replaces lfmm2d_ (src/laplace/lfmm2d.f:29-32; the C translation calls it
 * through cfmm2d, lfmm2d_ndiv.c:166) */
void lfmm2d_(const fmm2d_int *nd, const double *eps, const fmm2d_int *ns,
             const double *sources, const fmm2d_int *ifcharge, const fmm2d_cplx *charge,
             const fmm2d_int *ifdipole, const fmm2d_cplx *dipstr, const double *dipvec,
             fmm2d_int *iper, const fmm2d_int *ifpgh, fmm2d_cplx *pot, fmm2d_cplx *grad,
             fmm2d_cplx *hess, const fmm2d_int *nt, const double *targ,
             const fmm2d_int *ifpghtarg, fmm2d_cplx *pottarg, fmm2d_cplx *gradtarg,
             fmm2d_cplx *hesstarg, fmm2d_int *ier);

/* replaces lfmm2d_ndiv_ (c_translation/src/lfmm2d_ndiv.c:57) */
void lfmm2d_ndiv_(const fmm2d_int *nd, const double *eps, const fmm2d_int *ns,
                  const double *sources, const fmm2d_int *ifcharge, const fmm2d_cplx *charge,
                  const fmm2d_int *ifdipole, const fmm2d_cplx *dipstr, const double *dipvec,
                  fmm2d_int *iper, const fmm2d_int *ifpgh, fmm2d_cplx *pot, fmm2d_cplx *grad,
                  fmm2d_cplx *hess, const fmm2d_int *nt, const double *targ,
                  const fmm2d_int *ifpghtarg, fmm2d_cplx *pottarg, fmm2d_cplx *gradtarg,
                  fmm2d_cplx *hesstarg, const fmm2d_int *ndiv, const fmm2d_int *idivflag,
                  const fmm2d_int *ifnear, double *timeinfo, fmm2d_int *ier);

/* ---- Laplace, log kernel with real densities: pot (nd,n), grad (nd,2,n),
 * hess (nd,3,n), all float64; charge/dipstr (nd,ns), dipvec (nd,2,ns) float64 */

/* replaces rfmm2d_ndiv_ (c_translation/src/rfmm2d_ndiv.c:51): the Cauchy
 * engine on charge and dipstr * (-(v1 + i v2)), outputs Re, (Re, -Im) and
 * (Re, -Im, -Re) of pot, d/dz and d2/dz2 (rfmm2d_ndiv.c:130-214) */
void rfmm2d_ndiv_(const fmm2d_int *nd, const double *eps, const fmm2d_int *ns,
                  const double *sources, const fmm2d_int *ifcharge, const double *charge,
                  const fmm2d_int *ifdipole, const double *dipstr, const double *dipvec,
                  fmm2d_int *iper, const fmm2d_int *ifpgh, double *pot, double *grad,
                  double *hess, const fmm2d_int *nt, const double *targ,
                  const fmm2d_int *ifpghtarg, double *pottarg, double *gradtarg,
                  double *hesstarg, const fmm2d_int *ndiv, const fmm2d_int *idivflag,
                  const fmm2d_int *ifnear, double *timeinfo, fmm2d_int *ier);

/* replaces the Fortran rfmm2d (src/laplace/rfmm2d.f:27-30; ndiv by the eps
 * table of lndiv2d, as it reaches it through cfmm2d, rfmm2d.f:155) */
void rfmm2d_(const fmm2d_int *nd, const double *eps, const fmm2d_int *ns, const double *sources,
             const fmm2d_int *ifcharge, const double *charge, const fmm2d_int *ifdipole,
             const double *dipstr, const double *dipvec, fmm2d_int *iper,
             const fmm2d_int *ifpgh, double *pot, double *grad, double *hess,
             const fmm2d_int *nt, const double *targ, const fmm2d_int *ifpghtarg,
             double *pottarg, double *gradtarg, double *hesstarg, fmm2d_int *ier);

/* ---- Helmholtz: pot = (i/4) sum c H0(k r) + dipoles, grad (nd,2,n), hess (nd,3,n) */

/* replaces hfmm2d_ (c_translation/src/hfmm2d.c:1529) */
void hfmm2d_(const fmm2d_int *nd, const double *eps, const fmm2d_cplx *zk, const fmm2d_int *ns,
             const double *sources, const fmm2d_int *ifcharge, const fmm2d_cplx *charge,
             const fmm2d_int *ifdipole, const fmm2d_cplx *dipstr, const double *dipvec,
             fmm2d_int *iper, const fmm2d_int *ifpgh, fmm2d_cplx *pot, fmm2d_cplx *grad,
             fmm2d_cplx *hess, const fmm2d_int *nt, const double *targ,
             const fmm2d_int *ifpghtarg, fmm2d_cplx *pottarg, fmm2d_cplx *gradtarg,
             fmm2d_cplx *hesstarg, fmm2d_int *ier);

/* replaces hfmm2d_ndiv_ (c_translation/src/hfmm2d_ndiv.c:68) */
void hfmm2d_ndiv_(const fmm2d_int *nd, const double *eps, const fmm2d_cplx *zk,
                  const fmm2d_int *ns, const double *sources, const fmm2d_int *ifcharge,
                  const fmm2d_cplx *charge, const fmm2d_int *ifdipole, const fmm2d_cplx *dipstr,
                  const double *dipvec, fmm2d_int *iper, const fmm2d_int *ifpgh,
                  fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const fmm2d_int *nt,
                  const double *targ, const fmm2d_int *ifpghtarg, fmm2d_cplx *pottarg,
                  fmm2d_cplx *gradtarg, fmm2d_cplx *hesstarg, const fmm2d_int *ndiv,
                  const fmm2d_int *idivflag, const fmm2d_int *ifnear, double *timeinfo,
                  fmm2d_int *ier);

/* ---- parameter helpers */

/* replaces h2dterms_ (c_translation/src/h2dterms.c:22) */
void h2dterms_(const double *bsize, const fmm2d_cplx *zk, const double *eps, fmm2d_int *nterms,
               fmm2d_int *ier);

/* replaces hndiv2d_ (c_translation/src/hndiv2d.c:15) */
void hndiv2d_(const double *eps, const fmm2d_int *ns, const fmm2d_int *nt,
              const fmm2d_int *ifcharge, const fmm2d_int *ifdipole, const fmm2d_int *ifpgh,
              const fmm2d_int *ifpghtarg, fmm2d_int *ndiv, fmm2d_int *idivflag);
/* replaces lndiv2d_ (src/laplace/lndiv2d.f:1; never translated to C) */
void lndiv2d_(const double *eps, const fmm2d_int *ns, const fmm2d_int *nt,
              const fmm2d_int *ifcharge, const fmm2d_int *ifdipole, const fmm2d_int *ifpgh,
              const fmm2d_int *ifpghtarg, fmm2d_int *ndiv, fmm2d_int *idivflag);
/* replaces l2dterms_ (c_translation/src/l2dterms.c:26) */
void l2dterms_(const double *eps, fmm2d_int *nterms, fmm2d_int *ier);

/* ---- direct sums on the GPU (the MEX c2ddir oracle, matlab/fmm2d.mw:1426-1457).
 * Each replaces the routine of the same name in
 * c_translation/src/cauchykernels2d.c:31-442: outputs are INCREMENTED,
 * pairs closer than thresh are skipped with the routine's own comparison. */
void c2d_directcp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                   const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt,
                   fmm2d_cplx *pot, const double *thresh);
void c2d_directcg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                   const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt,
                   fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh);
void c2d_directch_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                   const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt,
                   fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const double *thresh);
void c2d_directdp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                   const fmm2d_cplx *dipstr, const double *targ, const fmm2d_int *nt,
                   fmm2d_cplx *pot, const double *thresh);
void c2d_directdg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                   const fmm2d_cplx *dipstr, const double *targ, const fmm2d_int *nt,
                   fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh);
void c2d_directdh_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                   const fmm2d_cplx *dipstr, const double *targ, const fmm2d_int *nt,
                   fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const double *thresh);
void c2d_directcdp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                    const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *targ,
                    const fmm2d_int *nt, fmm2d_cplx *pot, const double *thresh);
void c2d_directcdg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                    const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *targ,
                    const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad,
                    const double *thresh);
void c2d_directcdh_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                    const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *targ,
                    const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess,
                    const double *thresh);


/* ---- direct sums, log kernel with complex densities: replace l2d_direct{c,d,cd}{p,g,h}_
 * (c_translation/src/lapkernels2d.c:19-494); outputs incremented, grad (nd,2,nt),
 * hess (nd,3,nt) */
void l2d_directcp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, const double *thresh);
void l2d_directcg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh);
void l2d_directch_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const double *thresh);
void l2d_directdp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, const double *thresh);
void l2d_directdg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh);
void l2d_directdh_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const double *thresh);
void l2d_directcdp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, const double *thresh);
void l2d_directcdg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh);
void l2d_directcdh_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const double *thresh);

/* ---- direct sums, log kernel with real densities: replace r2d_direct{c,d,cd}{p,g,h}_
 * (c_translation/src/rlapkernels2d.c:16-491); outputs incremented, all float64,
 * grad (nd,2,nt), hess (nd,3,nt) */
void r2d_directcp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *charge, const double *targ, const fmm2d_int *nt, double *pot, const double *thresh);
void r2d_directcg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *charge, const double *targ, const fmm2d_int *nt, double *pot, double *grad, const double *thresh);
void r2d_directch_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *charge, const double *targ, const fmm2d_int *nt, double *pot, double *grad, double *hess, const double *thresh);
void r2d_directdp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, double *pot, const double *thresh);
void r2d_directdg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, double *pot, double *grad, const double *thresh);
void r2d_directdh_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, double *pot, double *grad, double *hess, const double *thresh);
void r2d_directcdp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *charge, const double *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, double *pot, const double *thresh);
void r2d_directcdg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *charge, const double *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, double *pot, double *grad, const double *thresh);
void r2d_directcdh_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns, const double *charge, const double *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, double *pot, double *grad, double *hess, const double *thresh);

/* ---- direct sums, Helmholtz (i/4) H0(k r): replace h2d_direct{c,d,cd}{p,g,h}_
 * (c_translation/src/helmkernels2d.c:21-520); outputs incremented */
void h2d_directcp_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, const double *thresh);
void h2d_directcg_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh);
void h2d_directch_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const double *thresh);
void h2d_directdp_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, const double *thresh);
void h2d_directdg_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh);
void h2d_directdh_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const double *thresh);
void h2d_directcdp_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, const double *thresh);
void h2d_directcdg_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh);
void h2d_directcdh_(const fmm2d_int *nd, const fmm2d_cplx *zk, const double *sources, const fmm2d_int *ns, const fmm2d_cplx *charge, const fmm2d_cplx *dipstr, const double *dipvec, const double *targ, const fmm2d_int *nt, fmm2d_cplx *pot, fmm2d_cplx *grad, fmm2d_cplx *hess, const double *thresh);

/* ---- biharmonic FMM: replaces bhfmm2d_ (c_translation/src/bhfmm2d.c:924,
 * include/bhfmm2d.h:22-31).  charge (nd,2,ns) and dip (nd,3,ns) complex;
 * pot (nd,n) and grad (nd,3,n) complex in the reference's layouts; leaf
 * size from bhndiv2d, order from bh2dterms.  As in the reference, pg 3
 * evaluates nothing (pot, grad and hess are written as zeros). */
void bhfmm2d_(const fmm2d_int *nd, const double *eps, const fmm2d_int *ns, const double *sources,
              const fmm2d_int *ifcharge, const fmm2d_cplx *charge, const fmm2d_int *ifdipole,
              const fmm2d_cplx *dip, fmm2d_int *iper, const fmm2d_int *ifpgh, fmm2d_cplx *pot,
              fmm2d_cplx *grad, fmm2d_cplx *hess, const fmm2d_int *nt, const double *targ,
              const fmm2d_int *ifpghtarg, fmm2d_cplx *pottarg, fmm2d_cplx *gradtarg,
              fmm2d_cplx *hesstarg, fmm2d_int *ier);
/* replaces stfmm2d_ (stfmm2d.c:40, include/stfmm2d.h:23-32): Stokeslets and
 * stresslets (nd,2,ns) to velocity (nd,2,n), pressure (nd,n) and velocity
 * gradient (nd,2,2,n) through the biharmonic FMM; ifppreg 1/2/3. */
void stfmm2d_(const fmm2d_int *nd, const double *eps, const fmm2d_int *nsource,
              const double *source, const fmm2d_int *ifstoklet, const double *stoklet,
              const fmm2d_int *ifstrslet, const double *strslet, const double *strsvec,
              const fmm2d_int *ifppreg, double *pot, double *pre, double *grad,
              const fmm2d_int *ntarg, const double *targ, const fmm2d_int *ifppregtarg,
              double *pottarg, double *pretarg, double *gradtarg, fmm2d_int *ier);
/* replace bh2dterms_ (bh2dterms.c:32) and bhndiv2d_ (bhndiv2d.c:17) */
void bh2dterms_(const double *eps, fmm2d_int *nterms, fmm2d_int *ier);
void bhndiv2d_(const double *eps, const fmm2d_int *ns, const fmm2d_int *nt,
               const fmm2d_int *ifcharge, const fmm2d_int *ifdipole, const fmm2d_int *ifpgh,
               const fmm2d_int *ifpghtarg, fmm2d_int *ndiv, fmm2d_int *idivflag);
/* replace bh2d_direct{c,d,cd}{p,g}_ (bhkernels2d.c:30-398, include/
 * bhkernels2d.h:34-67): vel (nd,nt) and grad (nd,3,nt) incremented, pairs
 * with |t - s| <= thresh skipped. */
void bh2d_directcp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                    const fmm2d_cplx *charges, const double *targ, const fmm2d_int *nt,
                    fmm2d_cplx *vel, const double *thresh);
void bh2d_directcg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                    const fmm2d_cplx *charges, const double *targ, const fmm2d_int *nt,
                    fmm2d_cplx *vel, fmm2d_cplx *grad, const double *thresh);
void bh2d_directdp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                    const fmm2d_cplx *dippar, const double *targ, const fmm2d_int *nt,
                    fmm2d_cplx *vel, const double *thresh);
void bh2d_directdg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                    const fmm2d_cplx *dippar, const double *targ, const fmm2d_int *nt,
                    fmm2d_cplx *vel, fmm2d_cplx *grad, const double *thresh);
void bh2d_directcdp_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                     const fmm2d_cplx *charges, const fmm2d_cplx *dippar, const double *targ,
                     const fmm2d_int *nt, fmm2d_cplx *vel, const double *thresh);
void bh2d_directcdg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ns,
                     const fmm2d_cplx *charges, const fmm2d_cplx *dippar, const double *targ,
                     const fmm2d_int *nt, fmm2d_cplx *vel, fmm2d_cplx *grad,
                     const double *thresh);

/* ---- direct sums, Stokes: replace st2ddirectstokg_ and st2ddirectstokstrsg_
 * (c_translation/src/stokkernels2d.c:11-173, MEX-called through st2ddir,
 * matlab/fmm2d.mw): Stokeslets stoklet (nd,2,ns) and stresslets
 * strslet/strsvec (nd,2,ns) to velocity pot (nd,2,nt), pressure pre (nd,nt)
 * and velocity gradient grad (nd,2,2,nt), all float64 and incremented;
 * pairs with r^2 < thresh^2 skipped. */
void st2ddirectstokg_(const fmm2d_int *nd, const double *sources, const double *stoklet,
                      const fmm2d_int *ns, const double *targ, const fmm2d_int *nt, double *pot,
                      double *pre, double *grad, const double *thresh);
void st2ddirectstokstrsg_(const fmm2d_int *nd, const double *sources, const fmm2d_int *ifstoklet,
                          const double *stoklet, const fmm2d_int *istress, const double *strslet,
                          const double *strsvec, const fmm2d_int *ns, const double *targ,
                          const fmm2d_int *nt, double *pot, double *pre, double *grad,
                          const double *thresh);

/* ---- allocation-free direct sum.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Arguments and
 * result as h2d_directcdg_ (helmkernels2d.c), run on `stream` (a
 * hipStream_t; NULL for the null stream), with all device memory -- staged
 * copies of the host arrays and the per-chunk partial sums -- taken from a
 * fixed 64 MiB array the library holds on each GPU until the process ends;
 * no hipMalloc, hipFree or other allocating call is made.  Returns 0; 107
 * when the call needs more than that array (about 1e6 sources with 1e3
 * targets), the outputs then untouched; 101 / 104 for HIP / device errors.
 * Every call on one GPU shares its array: calls that overlap on one GPU
 * (from several threads or streams) race. */
int fmm2dhip_h2d_directcdg_noalloc(const fmm2d_int *nd, const fmm2d_cplx *zk,
                                   const double *sources, const fmm2d_int *ns,
                                   const fmm2d_cplx *charge, const fmm2d_cplx *dipstr,
                                   const double *dipvec, const double *targ, const fmm2d_int *nt,
                                   fmm2d_cplx *pot, fmm2d_cplx *grad, const double *thresh,
                                   void *stream);

/* ---- device-resident variants (no counterpart in the reference).
 * Same argument meaning as cfmm2d_ndiv_ / lfmm2d_ndiv_, but every array is a
 * device pointer, and `stream` (a hipStream_t, or NULL for the library's
 * own) orders the work.  For callers whose data already lives in HBM. */
void fmm2dhip_cfmm2d_dev(int nd, double eps, int ns, const double *sources, int ifcharge,
                         const void *charge, int ifdipole, const void *dipstr, int ifpgh,
                         void *pot, void *grad, void *hess, int nt, const double *targ,
                         int ifpghtarg, void *pottarg, void *gradtarg, void *hesstarg,
                         int ndiv, int idivflag, int ifnear, double *timeinfo, void *stream,
                         int *ier);
void fmm2dhip_lfmm2d_dev(int nd, double eps, int ns, const double *sources, int ifcharge,
                         const void *charge, int ifdipole, const void *dipstr,
                         const double *dipvec, int ifpgh, void *pot, void *grad, void *hess,
                         int nt, const double *targ, int ifpghtarg, void *pottarg,
                         void *gradtarg, void *hesstarg, int ndiv, int idivflag, int ifnear,
                         double *timeinfo, void *stream, int *ier);

void fmm2dhip_hfmm2d_dev(int nd, double eps, double zk_re, double zk_im, int ns,
                         const double *sources, int ifcharge, const void *charge, int ifdipole,
                         const void *dipstr, const double *dipvec, int ifpgh, void *pot,
                         void *grad, void *hess, int nt, const double *targ, int ifpghtarg,
                         void *pottarg, void *gradtarg, void *hesstarg, int ndiv, int idivflag,
                         int ifnear, double *timeinfo, void *stream, int *ier);

/* ---- test hook (no counterpart in the reference): builds the box-level
 * tree and lists on the host, from Morton keys computed with the engine's
 * key recurrence and sorted on the CPU, in the reference's own box numbering
 * and itree layout (pts_tree2d.c:484-491, 1-based ids), plus the point
 * permutations (1-based) and per-box point ranges, and lists as
 * nlist(4,nboxes) and lists(65,nboxes) = list1(13) list2(27) list3(20)
 * list4(5).  Call with itree == NULL to get the sizes.  The solvers build
 * their tree on the device (fmm2dhip_test_devtree); this host restatement
 * lets CPU-only tests pin the key recurrence and box arithmetic against the
 * reference's trees. */
int fmm2dhip_test_tree(const double *src, int ns, const double *targ, int nt, int idivflag,
                       int ndiv, int *nlevels, int *nboxes, int *itree, double *centers,
                       double *boxsize, int *isrc, int *isrcse, int *itarg, int *itargse,
                       int *nlist, int *lists);

/* ---- test hook (no counterpart in the reference): the front end's device
 * radix sort.  Sorts n keys of key_bytes (4 or 8) bytes stably by key bits
 * [begin, end): keys_out receives the sorted keys and, when vals_out is not
 * NULL, vals_out[i] the input position of the i-th sorted key.  Returns 0,
 * 102 for bad arguments, or 101 when the sort reports a failed wait. */
int fmm2dhip_test_sort(const void *keys, int n, int key_bytes, int begin, int end,
                       void *keys_out, int *vals_out);

/* ---- test hook (no counterpart in the reference): the diagonal-form
 * translations' FFT.  Transforms nrows rows of n complex values (interleaved
 * re, im; n an even product of powers of 2, 3 and 5, as next235 returns),
 * unnormalised: sign -1 sum_j x_j e^{-2 pi i jk/n} (FFTPACK zfftf), +1 with
 * e^{+...} (zfftb).  Returns 0, or 102 for bad arguments. */
int fmm2dhip_test_fft(int n, int nrows, int sign, const double *in, double *out);

/* ---- test hook (no counterpart in the reference): the solvers' device
 * tree for the given points (copied to the device first; pg = pgt = 1 work
 * lists).  Boxes are numbered level-major, Morton order inside a level
 * (0-based).  boxinfo(12,nboxes) = level, ix, iy (cell indices at the
 * box's level), nchild, sorted source range [s0, s1), target range [t0, t1),
 * n1..n4; centers(2,nboxes); lists = for each box its list 1, 2, 3 and 4
 * entries (box ids) back to back; isrc/itarg = original index of each sorted
 * position (0-based, original order inside each leaf).  Returns 102 with
 * nlevels/nboxes/nlists set when nboxes > cap_boxes or nlists > cap_lists. */
int fmm2dhip_test_devtree(const double *src, int ns, const double *targ, int nt, int idivflag,
                          int ndiv, int cap_boxes, int cap_lists, int *nlevels, int *nboxes,
                          int *nlists, int *boxinfo, double *centers, int *lists, int *isrc,
                          int *itarg);

/* work counts and step times of the calling thread's last solve (no
 * counterpart in the reference): out[0] list-1 pairs at sources, [1] at
 * targets, [2] list-2 pairs, [3] M2L flops (8 ncoef^2 per pair), [4] boxes,
 * [5] levels, [6] largest expansion order, [7..14] step times in seconds in
 * timeinfo order, [15..22] host wall times of the front end (bounding box,
 * keys and sort, adaptive refinement, level restriction, boxes + lists +
 * work lists, leaf order and gather, operator setup, whole call including
 * the end-of-solve synchronisation), [23] device workspace bytes the solve
 * took, [24] and [25] the near field's source reads: over the leaves
 * holding sources ([24]) and targets ([25]), the sum of the source counts
 * of their list 1.  Returns the number of values available (26). */
int fmm2dhip_last_stats(double *out, int n);

/* frees the calling thread's device workspace on the current device (no
 * counterpart in the reference).  Each thread keeps, per device, the device
 * memory of its largest solve so far and reuses it for the next solve;
 * this returns it to the system.  Safe to call between solves only. */
void fmm2dhip_release_workspace(void);

/* arithmetic of the list-1 (near field) and M2L steps for the calling
 * thread's later solves (no counterpart in the reference, which is FP64
 * throughout).  FMM2DHIP_PREC_FP64: every step in FP64 (the default, unless
 * the environment variable FMM2DHIP_PRECISION is "mixed" when the library
 * loads).  FMM2DHIP_PREC_MIXED: when eps >= 1e-6, the Laplace near field
 * and the M2L translations of both kernels compute in FP32 (differences of
 * coordinates in FP64, sums carried in FP64); tighter eps stays FP64.
 * Results then agree with the FP64 solve to about eps rather than to
 * rounding.  Returns the previous mode; other values leave it unchanged. */
#define FMM2DHIP_PREC_FP64 0
#define FMM2DHIP_PREC_MIXED 1
int fmm2dhip_set_precision(int mode);
int fmm2dhip_get_precision(void);

/* library/device information for diagnostics: returns 0 when a gfx950
 * device is usable and writes its name (at most len bytes). */
int fmm2dhip_device_info(char *name, int len);

/* ---- batched and served solves (no counterpart in the reference, whose
 * entry points solve one problem per call).
 *
 * A problem of any of the four kernels, by value.  Field meanings are those
 * of the drop-in calls (cfmm2d_ndiv_, lfmm2d_ndiv_, hfmm2d_ndiv_,
 * rfmm2d_ndiv_: real strengths and outputs for FMM2DHIP_REAL); arrays
 * are host memory, or device memory in the layouts of the _dev entry points
 * when device_ptrs is 1.  Outputs are overwritten as by those calls.  The
 * caller keeps the struct and its arrays alive until the problem is done. */
#define FMM2DHIP_CAUCHY 0    /* cfmm2d */
#define FMM2DHIP_LOG 1       /* lfmm2d */
#define FMM2DHIP_HELMHOLTZ 2 /* hfmm2d */
#define FMM2DHIP_REAL 3      /* rfmm2d */
typedef struct fmm2dhip_problem {
  int kernel;
  int nd;
  double eps;
  double zk_re, zk_im; /* Helmholtz wavenumber */
  int ns;
  const double *sources;
  int ifcharge;
  const void *charge;
  int ifdipole;
  const void *dipstr;
  const double *dipvec; /* log, real and Helmholtz kernels */
  int ifpgh;
  void *pot, *grad, *hess;
  int nt;
  const double *targ;
  int ifpghtarg;
  void *pottarg, *gradtarg, *hesstarg;
  int ndiv;        /* <= 0: the leaf-size rule for eps (hndiv2d_) */
  int precision;   /* FMM2DHIP_PREC_FP64 or FMM2DHIP_PREC_MIXED */
  int device_ptrs; /* 1: every array is a device pointer */
  /* written by the library */
  int ier;
  double seconds;  /* wall time of the solve on its worker */
  double latency;  /* submission to completion (s), queueing included */
} fmm2dhip_problem;

/* Solves probs[0..n) on nworkers worker threads (1..32), each with its own
 * HIP stream and resident device workspace, so the problems' kernels
 * overlap on the device; larger problems start first.  The workers persist
 * in the process and are reused by later calls.  Returns the number of
 * problems whose ier is not 0. */
int fmm2dhip_batch(int n, fmm2dhip_problem *probs, int nworkers);

/* Solves probs[0..n) fused: problems that share kernel, nd, eps, leaf size
 * (ndiv), charge/dipole flags, ifpgh/ifpghtarg, whether they have targets,
 * precision and device_ptrs, and have more sources than the leaf size (and,
 * for Helmholtz, a real positive zk), are solved together as one forest:
 * each problem's tree sits below its own cell of one quadtree, so the whole
 * group runs as one level-synchronous launch sequence on the calling
 * thread's stream (one point sort, one tree build, one launch per FMM step
 * and level).  Every other problem runs on the fmm2dhip_batch worker pool
 * meanwhile.  Laplace results (cfmm2d, lfmm2d, rfmm2d) equal the problem's
 * solve alone up to rounding (same tree, same expansions); a Helmholtz
 * problem is rescaled to the group's common wavenumber, its root box
 * enlarged at most twofold, and agrees with its solve alone within eps.
 * Outputs, ier and seconds (the group's wall time) are written as by
 * fmm2dhip_batch; returns the number of problems whose ier is not 0. */
int fmm2dhip_batch_fused(int n, fmm2dhip_problem *probs, int nworkers);

/* Plans: repeated solves on fixed points (iterative solvers, time steps).
 * fmm2dhip_plan_create takes problems whose arrays are all device memory
 * (device_ptrs = 1), solves them once as fmm2dhip_batch_fused would, then
 * builds each group's (or lone problem's) point sort, tree and lists once
 * and captures the rest of its solve -- strengths copied in, every FMM
 * step, outputs written -- as a HIP graph.  fmm2dhip_plan_execute replays
 * the graphs on the calling thread's stream: it reads the current contents
 * of each problem's charge, dipstr and dipvec arrays and overwrites its
 * outputs, with the results an fmm2dhip_batch_fused call would give on them;
 * the sources, targets, flags and array pointers are those of creation and
 * must stay valid and unchanged.  The plan holds its device memory until
 * fmm2dhip_plan_destroy.  create returns NULL and sets *ier (102 for host
 * arrays) on failure; execute returns 0 or an ier and writes each problem's
 * ier and seconds.  Executes of one plan must not overlap. */
typedef struct fmm2dhip_plan fmm2dhip_plan;
fmm2dhip_plan *fmm2dhip_plan_create(int n, fmm2dhip_problem *probs, int *ier);
int fmm2dhip_plan_execute(fmm2dhip_plan *plan);
void fmm2dhip_plan_destroy(fmm2dhip_plan *plan);

/* Autotuner of the translation kernels' tile shapes (no counterpart in the
 * reference).  The process-wide table holds, for each M2L order's count MT
 * of 16-row tiles (1..8), Laplace (with biharmonic) and Helmholtz, FP64 and
 * FP32 apart, the row tiles each wave of the M2L kernels computes
 * (1..min(4, MT)), and the level size (in 16-row wave tiles) below which
 * M2M/L2L split their slots over wave groups.  fmm2dhip_autotune solves *p
 * (any kernel, host or device arrays; outputs overwritten) 1 + nrep times
 * per candidate, times the M2L (and M2M + L2L) steps with the solve's own
 * step events, and keeps each choice that is at least 2% faster than the
 * current one; it returns the number of entries changed, or -ier.
 * fmm2dhip_autotune_batch does the same for every group of probs that
 * fmm2dhip_batch_fused would solve as one forest, timing the group's forest.
 * fmm2dhip_tune_get writes the table as 37 ints: 9 per (family, precision)
 * in the order Laplace FP64, Laplace FP32, Helmholtz FP64, Helmholtz FP32
 * (MT 0..8, index 0 unused), then the M2M/L2L threshold; save/load write and
 * read it as text (0 on success for save, the entries read for load, -1
 * when the file cannot be opened).  At start the table is loaded from the
 * file FMM2DHIP_TUNE_FILE names, if any. */
int fmm2dhip_autotune(fmm2dhip_problem *p, int nrep);
int fmm2dhip_autotune_batch(int n, fmm2dhip_problem *probs, int nrep);
int fmm2dhip_tune_get(int *out, int n);
int fmm2dhip_tune_save(const char *path);
int fmm2dhip_tune_load(const char *path);

/* A persistent request loop.  Requests are admitted in arrival order as a
 * worker and device memory free up: a worker keeps the workspace of its
 * largest solve resident, and a request starts only while the workspaces of
 * all workers, its own grown to the request's estimate, stay within
 * hbm_budget_bytes (idle workers give theirs back when the head of the
 * queue would not fit).  Work of concurrent requests overlaps on the
 * device.  submit returns a ticket >= 0, or -1 when the request alone
 * exceeds the budget or the server is shutting down; wait returns 0 when
 * the ticket is done, 1 on time-out, -1 for an unknown ticket; a ticket
 * is forgotten once a wait has returned 0 for it (a later wait returns -1).
 * stats writes up to n of: requests done, failed, points solved, seconds
 * from the first submission to the last completion, Mpts/s over that span,
 * p50, p99 and max latency (s), peak workspace bytes held, budget, tickets
 * still tracked (queued, running, or done and not yet waited for); returns
 * the count available.  destroy finishes admitted and queued work, then stops the
 * workers and frees their workspaces. */
typedef struct fmm2dhip_server fmm2dhip_server;
fmm2dhip_server *fmm2dhip_server_create(int nworkers, double hbm_budget_bytes);
long long fmm2dhip_server_submit(fmm2dhip_server *s, fmm2dhip_problem *p);
int fmm2dhip_server_wait(fmm2dhip_server *s, long long ticket, double timeout_s);
int fmm2dhip_server_stats(fmm2dhip_server *s, double *out, int n);
void fmm2dhip_server_destroy(fmm2dhip_server *s);

#ifdef __cplusplus
}
#endif
#endif /* FMM2D_HIP_H */
