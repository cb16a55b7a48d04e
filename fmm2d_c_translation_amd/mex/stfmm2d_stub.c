/*
 * stfmm2d_ for the MEX gateway: the gateway's stfmm2d stub
 * (matlab/fmm2d.mw, stfmm2d.m:74-75) calls the Stokes FMM
 * (c_translation/src/stfmm2d.c), which libfmm2d_hip does not provide.  This
 * definition raises a MEX error, so mex_dispatch returns 1 with the message
 * below (ier is set to 108 first) instead of the call returning zeros.  The Stokes direct sums
 * (st2ddirectstokg_, st2ddirectstokstrsg_, behind st2ddir.m) are in the
 * library and run on the GPU.
 */
void mexErrMsgTxt(const char *msg);

int stfmm2d_(int *nd, double *eps, int *ns, double *sources, int *ifstoklet, double *stoklet,
             int *ifstrslet, double *strslet, double *strsvec, int *ifppreg, double *pot,
             double *pre, double *grad, int *nt, double *targ, int *ifppregtarg, double *pottarg,
             double *pretarg, double *gradtarg, int *ier) {
    (void)nd, (void)eps, (void)ns, (void)sources, (void)ifstoklet, (void)stoklet, (void)ifstrslet,
        (void)strslet, (void)strsvec, (void)ifppreg, (void)pot, (void)pre, (void)grad, (void)nt,
        (void)targ, (void)ifppregtarg, (void)pottarg, (void)pretarg, (void)gradtarg;
    *ier = 108;
    mexErrMsgTxt("stfmm2d: the Stokes FMM is not provided by libfmm2d_hip "
                 "(the Stokes direct sums st2ddir are)");
    return 0;
}
