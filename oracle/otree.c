/*
 * otree.c -- oracle restatement of the adaptive level-restricted quadtree and
 * the FMM interaction lists (test infrastructure; see oracle.h).
 *
 * Follows, routine by routine:
 *   bounding box, refinement loop   c_translation/src/pts_tree2d.c:493-610
 *   child creation                  c_translation/src/tree_routs2d.c:48-101
 *   stable 4-way point partition    c_translation/src/pts_tree2d.c:670-761
 *   colleagues                      c_translation/src/tree_routs2d.c:140-185
 *   2:1 level restriction           c_translation/src/pts_tree2d.c:767-960
 *   flag-driven refinement          c_translation/src/tree_routs2d.c:252-318
 *   flag update                     c_translation/src/tree_routs2d.c:197-245
 *   renumbering after the fix       c_translation/src/pts_tree2d.c:966-1081
 *   final point sort                c_translation/src/pts_tree2d.c:1087-1119
 *   lists 1-4                       c_translation/src/tree_routs2d.c:369-510
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define LO(l) (2 * (l))
#define HI(l) (2 * (l) + 1)

static void *xrealloc(void *p, size_t n) { return realloc(p, n ? n : 1); }

/* make room for box ids up to `need` (inclusive) */
static void tree_reserve(otree *t, int need, int **iflag)
{
    if (need < t->cap) return;
    int nc = t->cap ? t->cap : 64;
    while (nc <= need) nc *= 2;
    t->ilevel = xrealloc(t->ilevel, sizeof(int) * nc);
    t->iparent = xrealloc(t->iparent, sizeof(int) * nc);
    t->nchild = xrealloc(t->nchild, sizeof(int) * nc);
    t->ichild = xrealloc(t->ichild, sizeof(int) * 4 * nc);
    t->nnbors = xrealloc(t->nnbors, sizeof(int) * nc);
    t->nbors = xrealloc(t->nbors, sizeof(int) * 9 * nc);
    t->centers = xrealloc(t->centers, sizeof(double) * 2 * nc);
    if (iflag) *iflag = xrealloc(*iflag, sizeof(int) * nc);
    t->cap = nc;
}

/* Appends four children for every box of [first,last] whose flag is set.
 * Children of the k-th flagged box get ids nb+4k+1..nb+4k+4 in the order
 * (-x,-y), (+x,-y), (-x,+y), (+x,+y); their centres are parent +- bs/2
 * (tree_routs2d.c:70-92).  With a non-NULL iflag the flag-driven variant
 * is used: children of flag-1 boxes become flag 3, of flag-2 boxes flag 0
 * (tree_routs2d.c:282-305). */
static void refine_range(otree *t, int first, int last, const char *want,
                         int **iflag, double bs, int child_level)
{
    int nb = t->nboxes, k = 0;
    for (int ib = first; ib <= last; ib++) {
        int on = want ? want[ib - first] : ((*iflag)[ib] > 0);
        if (!on) continue;
        int base = nb + 4 * k;
        k++;
        tree_reserve(t, base + 4, iflag);
        t->nchild[ib] = 4;
        for (int j = 1; j <= 4; j++) {
            int jb = base + j;
            double sx = (j % 2 == 0) ? 1.0 : -1.0;
            double sy = (j <= 2) ? -1.0 : 1.0;
            t->centers[2 * jb] = t->centers[2 * ib] + sx * bs / 2;
            t->centers[2 * jb + 1] = t->centers[2 * ib + 1] + sy * bs / 2;
            t->iparent[jb] = ib;
            t->nchild[jb] = 0;
            for (int l = 0; l < 4; l++) t->ichild[4 * jb + l] = -1;
            t->ichild[4 * ib + j - 1] = jb;
            t->ilevel[jb] = child_level;
            if (!want) {
                if ((*iflag)[ib] == 1) (*iflag)[jb] = 3;
                if ((*iflag)[ib] == 2) (*iflag)[jb] = 0;
            }
        }
    }
    t->nboxes = nb + 4 * k;
}

/* Stable split of the points of box ib into its four children: lower half
 * (y - cy < 0) before upper, then left (x - cx < 0) before right inside
 * each half, original order kept (pts_tree2d.c:694-757). */
static void split_points(const otree *t, int ib, const double *xy, int *perm,
                         int *se, int *tmp)
{
    int s = se[2 * ib], e = se[2 * ib + 1];
    double cx = t->centers[2 * ib], cy = t->centers[2 * ib + 1];
    int cnt[4] = {0, 0, 0, 0};
    for (int pass = 0; pass < 4; pass++) {
        for (int i = s; i <= e; i++) {
            int p = perm[i - 1];
            int q = (xy[2 * (p - 1)] - cx < 0 ? 0 : 1) +
                    (xy[2 * (p - 1) + 1] - cy < 0 ? 0 : 2);
            if (q == pass) tmp[cnt[0] + cnt[1] + cnt[2] + cnt[3]] = p, cnt[q]++;
        }
    }
    for (int i = s; i <= e; i++) perm[i - 1] = tmp[i - s];
    int start = s;
    for (int j = 0; j < 4; j++) {
        int jb = t->ichild[4 * ib + j];
        se[2 * jb] = start;
        se[2 * jb + 1] = start + cnt[j] - 1;
        start += cnt[j];
    }
}

static int touches(const otree *t, int a, int b, double d)
{
    return fabs(t->centers[2 * a] - t->centers[2 * b]) <= d &&
           fabs(t->centers[2 * a + 1] - t->centers[2 * b + 1]) <= d;
}

/* colleagues: children of the parent's colleagues within 1.05 bs
 * (tree_routs2d.c:163-184) */
static void colleagues_of(otree *t, int ib, double bs)
{
    int dad = t->iparent[ib];
    t->nnbors[ib] = 0;
    for (int i = 0; i < t->nnbors[dad]; i++) {
        int jb = t->nbors[9 * dad + i];
        for (int j = 0; j < 4; j++) {
            int kb = t->ichild[4 * jb + j];
            if (kb > 0 && touches(t, kb, ib, 1.05 * bs))
                t->nbors[9 * ib + t->nnbors[ib]++] = kb;
        }
    }
}

static void all_colleagues(otree *t)
{
    for (int i = 1; i <= t->nboxes; i++) {
        t->nnbors[i] = 0;
        for (int j = 0; j < 9; j++) t->nbors[9 * i + j] = -1;
    }
    t->nnbors[1] = 1;
    t->nbors[9 * 1] = 1;
    for (int l = 1; l <= t->nlevels; l++)
        for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++)
            colleagues_of(t, ib, t->boxsize[l]);
}

/* renumber so that every level lists its original boxes and then its tail
 * boxes contiguously (pts_tree2d.c:966-1081) */
static void reorg(otree *t, const int *tail, int *iflag)
{
    int n = t->nboxes;
    int *lev = malloc(sizeof(int) * (n + 1)), *par = malloc(sizeof(int) * (n + 1));
    int *nch = malloc(sizeof(int) * (n + 1)), *chi = malloc(sizeof(int) * 4 * (n + 1));
    int *flg = malloc(sizeof(int) * (n + 1)), *map = malloc(sizeof(int) * (n + 1));
    double *cen = malloc(sizeof(double) * 2 * (n + 1));
    int *tl = malloc(sizeof(int) * 2 * (t->nlevels + 1));
    memcpy(tl, t->laddr, sizeof(int) * 2 * (t->nlevels + 1));
    memcpy(lev, t->ilevel, sizeof(int) * (n + 1));
    memcpy(par, t->iparent, sizeof(int) * (n + 1));
    memcpy(nch, t->nchild, sizeof(int) * (n + 1));
    memcpy(chi, t->ichild, sizeof(int) * 4 * (n + 1));
    memcpy(flg, iflag, sizeof(int) * (n + 1));
    memcpy(cen, t->centers, sizeof(double) * 2 * (n + 1));
    for (int l = 0; l <= 1 && l <= t->nlevels; l++)
        for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) map[ib] = ib;
    if (t->nlevels >= 2) {
        int cur = t->laddr[LO(2)];
        for (int l = 2; l <= t->nlevels; l++) {
            t->laddr[LO(l)] = cur;
            for (int pass = 0; pass < 2; pass++) {
                int a = pass ? tail[LO(l)] : tl[LO(l)];
                int b = pass ? tail[HI(l)] : tl[HI(l)];
                for (int ib = a; ib <= b; ib++) {
                    t->ilevel[cur] = lev[ib];
                    t->nchild[cur] = nch[ib];
                    t->centers[2 * cur] = cen[2 * ib];
                    t->centers[2 * cur + 1] = cen[2 * ib + 1];
                    iflag[cur] = flg[ib];
                    map[ib] = cur++;
                }
            }
            t->laddr[HI(l)] = cur - 1;
        }
    }
    for (int ib = 1; ib <= n; ib++) {
        int m = map[ib];
        t->iparent[m] = par[ib] > 0 ? map[par[ib]] : -1;
        for (int j = 0; j < 4; j++) {
            int c = chi[4 * ib + j];
            t->ichild[4 * m + j] = c > 0 ? map[c] : -1;
        }
    }
    free(lev); free(par); free(nch); free(chi); free(flg); free(map); free(cen); free(tl);
}

/* flag++ boxes at `lev` stay refined only if a child of a colleague is
 * itself refined and touches them (tree_routs2d.c:214-244) */
static void update_flags(otree *t, int lev, int first, int last, int *iflag)
{
    double d = 1.05 * (t->boxsize[lev] + t->boxsize[lev + 1]) / 2.0;
    for (int ib = first; ib <= last; ib++) {
        if (iflag[ib] != 3) continue;
        iflag[ib] = 0;
        for (int i = 0; i < t->nnbors[ib] && !iflag[ib]; i++) {
            int jb = t->nbors[9 * ib + i];
            for (int j = 0; j < 4; j++) {
                int kb = t->ichild[4 * jb + j];
                if (kb > 0 && t->nchild[kb] > 0 && touches(t, kb, ib, d)) {
                    iflag[ib] = 1;
                    break;
                }
            }
        }
    }
}

/* 2:1 balance (pts_tree2d.c:767-960) */
static void fix_level_restriction(otree *t)
{
    int L = t->nlevels;
    int *iflag = calloc((size_t)t->cap, sizeof(int));
    int *tail = malloc(sizeof(int) * 2 * (L + 1));
    const double *bs = t->boxsize;

    /* flag 1: leaves touching a parent whose children exist two levels down */
    for (int l = L; l >= 2; l--) {
        double d = 1.05 * (bs[l - 1] + bs[l - 2]) / 2.0;
        for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
            int dad = t->iparent[ib], gd = t->iparent[dad];
            for (int i = 0; i < t->nnbors[gd]; i++) {
                int jb = t->nbors[9 * gd + i];
                if (t->nchild[jb] == 0 && iflag[jb] == 0 && touches(t, jb, dad, d))
                    iflag[jb] = 1;
            }
        }
    }
    /* flag 2: leaves touching a flagged box one level down */
    for (int l = L; l >= 1; l--) {
        double d = 1.05 * (bs[l] + bs[l - 1]) / 2.0;
        for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++) {
            if (iflag[ib] != 1 && iflag[ib] != 2) continue;
            int dad = t->iparent[ib];
            for (int i = 0; i < t->nnbors[dad]; i++) {
                int jb = t->nbors[9 * dad + i];
                if (t->nchild[jb] == 0 && iflag[jb] == 0 && touches(t, jb, ib, d))
                    iflag[jb] = 2;
            }
        }
    }
    for (int l = 0; l <= L; l++) tail[LO(l)] = 0, tail[HI(l)] = -1;
    for (int l = 1; l <= L - 2; l++) {
        tail[LO(l + 1)] = t->nboxes + 1;
        refine_range(t, t->laddr[LO(l)], t->laddr[HI(l)], NULL, &iflag, bs[l + 1], l + 1);
        tail[HI(l + 1)] = t->nboxes;
    }
    reorg(t, tail, iflag);
    all_colleagues(t);
    for (int ib = 1; ib <= t->nboxes; ib++)
        if (iflag[ib] != 3) iflag[ib] = 0;
    for (int l = 0; l <= L; l++) tail[LO(l)] = 0, tail[HI(l)] = -1;
    for (int l = 2; l <= L - 2; l++) {
        update_flags(t, l, t->laddr[LO(l)], t->laddr[HI(l)], iflag);
        update_flags(t, l, tail[LO(l)], tail[HI(l)], iflag);
        tail[LO(l + 1)] = t->nboxes + 1;
        refine_range(t, t->laddr[LO(l)], t->laddr[HI(l)], NULL, &iflag, bs[l + 1], l + 1);
        refine_range(t, tail[LO(l)], tail[HI(l)], NULL, &iflag, bs[l + 1], l + 1);
        tail[HI(l + 1)] = t->nboxes;
        for (int ib = tail[LO(l + 1)]; ib <= tail[HI(l + 1)]; ib++)
            colleagues_of(t, ib, bs[l + 1]);
    }
    reorg(t, tail, iflag);
    all_colleagues(t);
    free(iflag);
    free(tail);
}

/* final point order: split every parent level by level (pts_tree2d.c:1087-1119) */
static void sort_points(const otree *t, const double *xy, int n, int **perm, int **se)
{
    *perm = malloc(sizeof(int) * (n > 0 ? n : 1));
    *se = malloc(sizeof(int) * 2 * (t->nboxes + 1));
    int *tmp = malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) (*perm)[i] = i + 1;
    (*se)[2] = 1;
    (*se)[3] = n;
    for (int l = 0; l <= t->nlevels - 1; l++)
        for (int ib = t->laddr[LO(l)]; ib <= t->laddr[HI(l)]; ib++)
            if (t->nchild[ib] > 0) split_points(t, ib, xy, *perm, *se, tmp);
    free(tmp);
}

void otree_build(const double *src, int ns, const double *targ, int nt,
                 int idivflag, int ndiv, int nlmin, int nlmax, int ifunif,
                 otree *t)
{
    memset(t, 0, sizeof(*t));
    double xmin = src[0], xmax = src[0], ymin = src[1], ymax = src[1];
    for (int i = 0; i < ns; i++) {
        double x = src[2 * i], y = src[2 * i + 1];
        if (x < xmin) xmin = x;
        if (x > xmax) xmax = x;
        if (y < ymin) ymin = y;
        if (y > ymax) ymax = y;
    }
    for (int i = 0; i < nt; i++) {
        double x = targ[2 * i], y = targ[2 * i + 1];
        if (x < xmin) xmin = x;
        if (x > xmax) xmax = x;
        if (y < ymin) ymin = y;
        if (y > ymax) ymax = y;
    }
    t->boxsize = malloc(sizeof(double) * (nlmax + 1));
    t->laddr = malloc(sizeof(int) * 2 * (nlmax + 1));
    t->boxsize[0] = xmax - xmin;
    if (ymax - ymin > t->boxsize[0]) t->boxsize[0] = ymax - ymin;
    tree_reserve(t, 1, NULL);
    t->nboxes = 1;
    t->laddr[0] = t->laddr[1] = 1;
    t->ilevel[1] = 0;
    t->iparent[1] = -1;
    t->nchild[1] = 0;
    for (int j = 0; j < 4; j++) t->ichild[4 + j] = -1;
    t->centers[2] = (xmin + xmax) / 2;
    t->centers[3] = (ymin + ymax) / 2;

    /* working partition during refinement */
    int *ps = malloc(sizeof(int) * (ns > 0 ? ns : 1)), *pt = malloc(sizeof(int) * (nt > 0 ? nt : 1));
    int secap = 64;
    int *ss = malloc(sizeof(int) * 2 * secap), *ts = malloc(sizeof(int) * 2 * secap);
    int *tmp = malloc(sizeof(int) * ((ns > nt ? ns : nt) + 1));
    for (int i = 0; i < ns; i++) ps[i] = i + 1;
    for (int i = 0; i < nt; i++) pt[i] = i + 1;
    ss[2] = 1; ss[3] = ns; ts[2] = 1; ts[3] = nt;

    int lev;
    for (lev = 0; lev < nlmax; lev++) {
        int first = t->laddr[LO(lev)], last = t->laddr[HI(lev)], nb = last - first + 1;
        char *want = malloc((size_t)nb);
        int any = 0;
        for (int i = 0; i < nb; i++) {
            int ib = first + i, nn = 0;
            int nss = ss[2 * ib + 1] - ss[2 * ib] + 1, ntt = ts[2 * ib + 1] - ts[2 * ib] + 1;
            if (idivflag == 0) nn = nss;
            if (idivflag == 1) nn = ntt;
            if (idivflag == 2) nn = ntt > nss ? ntt : nss;
            want[i] = lev >= nlmin ? (nn > ndiv) : 1;
            any |= want[i];
        }
        if (ifunif == 1 && lev >= nlmin)
            for (int i = 0; i < nb; i++) want[i] = (char)any;
        if (!any) { free(want); break; }
        t->boxsize[lev + 1] = t->boxsize[lev] / 2;
        t->laddr[LO(lev + 1)] = t->nboxes + 1;
        refine_range(t, first, last, want, NULL, t->boxsize[lev + 1], lev + 1);
        t->laddr[HI(lev + 1)] = t->nboxes;
        if (t->nboxes + 1 > secap) {
            while (t->nboxes + 1 > secap) secap *= 2;
            ss = realloc(ss, sizeof(int) * 2 * secap);
            ts = realloc(ts, sizeof(int) * 2 * secap);
        }
        for (int i = 0; i < nb; i++) {
            if (!want[i]) continue;
            split_points(t, first + i, src, ps, ss, tmp);
            split_points(t, first + i, targ, pt, ts, tmp);
        }
        free(want);
    }
    t->nlevels = lev;
    free(ps); free(pt); free(ss); free(ts); free(tmp);

    all_colleagues(t);
    if (t->nlevels >= 2) fix_level_restriction(t);
    sort_points(t, src, ns, &t->isrc, &t->isrcse);
    sort_points(t, targ, nt, &t->itarg, &t->itargse);
}

void otree_free(otree *t)
{
    free(t->laddr); free(t->ilevel); free(t->iparent); free(t->nchild);
    free(t->ichild); free(t->nnbors); free(t->nbors); free(t->centers);
    free(t->boxsize); free(t->isrc); free(t->isrcse); free(t->itarg); free(t->itargse);
    memset(t, 0, sizeof(*t));
}

/* lists 1-4 (tree_routs2d.c:369-510).  The loop order matters: list1 and
 * list4 entries that a coarse leaf adds to a finer box land before that
 * box's own entries because levels are visited coarse to fine. */
void olists_build(const otree *t, olists *L)
{
    int n = t->nboxes;
    L->mn1 = 13; L->mn2 = 27; L->mn3 = 20; L->mn4 = 5;
    L->n1 = calloc(n + 1, sizeof(int)); L->l1 = calloc((size_t)L->mn1 * (n + 1), sizeof(int));
    L->n2 = calloc(n + 1, sizeof(int)); L->l2 = calloc((size_t)L->mn2 * (n + 1), sizeof(int));
    L->n3 = calloc(n + 1, sizeof(int)); L->l3 = calloc((size_t)L->mn3 * (n + 1), sizeof(int));
    L->n4 = calloc(n + 1, sizeof(int)); L->l4 = calloc((size_t)L->mn4 * (n + 1), sizeof(int));
    if (t->nchild[1] == 0) { L->n1[1] = 1; L->l1[L->mn1 * 1] = 1; }
    for (int lev = 1; lev <= t->nlevels; lev++) {
        double bs = t->boxsize[lev];
        for (int ib = t->laddr[LO(lev)]; ib <= t->laddr[HI(lev)]; ib++) {
            int dad = t->iparent[ib];
            for (int i = 0; i < t->nnbors[dad]; i++) {
                int jb = t->nbors[9 * dad + i];
                for (int j = 0; j < t->nchild[jb]; j++) {
                    int kb = t->ichild[4 * jb + j];
                    if (fabs(t->centers[2 * kb] - t->centers[2 * ib]) >= 1.05 * bs ||
                        fabs(t->centers[2 * kb + 1] - t->centers[2 * ib + 1]) >= 1.05 * bs)
                        L->l2[L->mn2 * ib + L->n2[ib]++] = kb;
                }
            }
            if (t->nchild[ib] != 0) continue;
            for (int i = 0; i < t->nnbors[ib]; i++) {
                int jb = t->nbors[9 * ib + i];
                if (t->nchild[jb] == 0) {
                    L->l1[L->mn1 * ib + L->n1[ib]++] = jb;
                    continue;
                }
                double d = 1.05 * (bs + t->boxsize[lev + 1]) / 2.0;
                for (int j = 0; j < t->nchild[jb]; j++) {
                    int kb = t->ichild[4 * jb + j];
                    double xd = fabs(t->centers[2 * kb] - t->centers[2 * ib]);
                    double yd = fabs(t->centers[2 * kb + 1] - t->centers[2 * ib + 1]);
                    if (xd < d && yd < d) {
                        L->l1[L->mn1 * ib + L->n1[ib]++] = kb;
                        L->l1[L->mn1 * kb + L->n1[kb]++] = ib;
                    } else {
                        L->l3[L->mn3 * ib + L->n3[ib]++] = kb;
                        L->l4[L->mn4 * kb + L->n4[kb]++] = ib;
                    }
                }
            }
        }
    }
}

void olists_free(olists *L)
{
    free(L->n1); free(L->l1); free(L->n2); free(L->l2);
    free(L->n3); free(L->l3); free(L->n4); free(L->l4);
}

/* ---- flat export for the tests: the reference itree layout ------------
 * itree = laddr(2,0:nlevels), ilevel, iparent, nchild, ichild(4,.),
 * nnbors, nbors(9,.)  (pts_tree2d.c:484-491). */
void *oracle_tree_new(const double *src, int ns, const double *targ, int nt,
                      int idivflag, int ndiv, int *nlevels, int *nboxes)
{
    otree *t = malloc(sizeof(otree));
    otree_build(src, ns, targ, nt, idivflag, ndiv, 0, 51, 0, t);
    *nlevels = t->nlevels;
    *nboxes = t->nboxes;
    return t;
}

void oracle_tree_export(void *h, int *itree, double *centers, double *boxsize,
                        int *isrc, int *isrcse, int *itarg, int *itargse)
{
    otree *t = h;
    int n = t->nboxes, L = t->nlevels, k = 0;
    for (int l = 0; l <= L; l++) itree[k++] = t->laddr[LO(l)], itree[k++] = t->laddr[HI(l)];
    for (int i = 1; i <= n; i++) itree[k++] = t->ilevel[i];
    for (int i = 1; i <= n; i++) itree[k++] = t->iparent[i];
    for (int i = 1; i <= n; i++) itree[k++] = t->nchild[i];
    for (int i = 1; i <= n; i++) for (int j = 0; j < 4; j++) itree[k++] = t->ichild[4 * i + j];
    for (int i = 1; i <= n; i++) itree[k++] = t->nnbors[i];
    for (int i = 1; i <= n; i++) for (int j = 0; j < 9; j++) itree[k++] = t->nbors[9 * i + j];
    memcpy(centers, t->centers + 2, sizeof(double) * 2 * n);
    memcpy(boxsize, t->boxsize, sizeof(double) * (L + 1));
    int ns = t->isrcse[3], nt = t->itargse[3];
    memcpy(isrc, t->isrc, sizeof(int) * ns);
    memcpy(isrcse, t->isrcse + 2, sizeof(int) * 2 * n);
    memcpy(itarg, t->itarg, sizeof(int) * nt);
    memcpy(itargse, t->itargse + 2, sizeof(int) * 2 * n);
}

void oracle_lists_export(void *h, int *nlist, int *lists)
{
    otree *t = h;
    olists L;
    olists_build(t, &L);
    int n = t->nboxes;
    for (int i = 1; i <= n; i++) {
        nlist[4 * (i - 1) + 0] = L.n1[i];
        nlist[4 * (i - 1) + 1] = L.n2[i];
        nlist[4 * (i - 1) + 2] = L.n3[i];
        nlist[4 * (i - 1) + 3] = L.n4[i];
        int *o = lists + 65 * (i - 1);
        for (int j = 0; j < 13; j++) o[j] = j < L.n1[i] ? L.l1[13 * i + j] : 0;
        for (int j = 0; j < 27; j++) o[13 + j] = j < L.n2[i] ? L.l2[27 * i + j] : 0;
        for (int j = 0; j < 20; j++) o[40 + j] = j < L.n3[i] ? L.l3[20 * i + j] : 0;
        for (int j = 0; j < 5; j++) o[60 + j] = j < L.n4[i] ? L.l4[5 * i + j] : 0;
    }
    olists_free(&L);
}

void oracle_tree_free(void *h)
{
    otree_free(h);
    free(h);
}
