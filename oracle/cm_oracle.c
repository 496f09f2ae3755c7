/*
 * cm_oracle.c -- CPU FP64 restatement of the ComplexMixtures.jl MDDF hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or calls this file:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
 * and there only as the checker (or the timed CPU baseline), never as the
 * thing measured or shipped.
 *
 * What it restates (reference files, relative to the ComplexMixtures.jl tree):
 *   mddf_frame!                 src/mddf.jl:361-429
 *   coordination_number_frame!  src/mddf.jl:438-471
 *   inbulk                      src/mddf.jl:55-57
 *   randomize_solvent!          src/mddf.jl:65-88
 *   MinimumDistance/update_md   src/minimum_distances.jl:12-39
 *   update_list! (self/cross)   src/minimum_distances.jl:81-120
 *   update_counters!            src/update_counters.jl:43-88
 *   update_group_count!         src/update_counters.jl:22-36
 *   atom_type / mol_index       src/update_counters.jl:9, src/minimum_distances.jl:72
 *   eulermat / move! / random_move!   src/rigid_body.jl:45-57, 73-80, 107-137
 *   setbin                      src/results.jl:28
 *   sphericalshellvolume / shellradius   src/results.jl:254-257, 272-275
 *   sum!                        src/results.jl:629-649
 *   finalresults!               src/results.jl:311-469
 *   parallel_setup (chunks)     src/parallel_setup.jl:7-57, src/mddf.jl:285-338
 *   cell_volume                 src/mddf.jl:350-352
 *   units.Angs3tocm3permol      src/io.jl:6-14
 *
 * Definitions that the reference leaves to third-party code, fixed here and
 * shared (as a definition, not as code) with the GPU library:
 *   - Minimum image: n = rint(Hinv * dr), dr' = dr - H n (fma chain), d2 = fma
 *     chain.  For pairs within the cutoff (cutoff < half the smallest plane
 *     spacing) this is the exact minimum image.  CellListMap computes the same
 *     distance from wrapped coordinates; its last-bit values are not
 *     reproducible outside Julia.
 *   - Within cutoff: d2 <= cutoff^2 (equality handling in CellListMap is not
 *     pinned by any reference test).
 *   - Ties in update_md: the reference keeps the later-visited pair, and the
 *     visit order is internal to CellListMap.  Here ties on d go to the
 *     lexicographically smallest (i, j).
 *   - Random numbers: Philox4x32-10 (Salmon et al., SC'11) keyed by the seed
 *     and counted by (slot, sample, frame, stream tag), replacing the
 *     Xoroshiro128+/StableRNG streams (not reproducible outside Julia).
 *   - Random centre: uniform fractional coordinates in the unit cell (the
 *     reference draws in a 10^4-times larger computing box and wraps, which is
 *     the same distribution).
 *   - sin/cos(2*pi*u): exact quarter-turn reduction plus Taylor polynomials
 *     in Horner form (coefficients (2pi)^n/n!), so that host and device give
 *     the same bits.
 *
 * Layouts: coordinates [frame][atom][3] doubles; unit cells [frame][9] column
 * major (first three values = lattice vector a); group counts group-major
 * [group][bin], like Result.*_group_count (Vector{Vector{Float64}}).
 * All indices are 0-based here (the reference is 1-based).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

#define ORC_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Philox4x32-10                                                             */
/* ------------------------------------------------------------------------ */
ORC_API void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* stream tags (4th counter word) */
#define TAG_REF 0x52454631u   /* reference-solute draws */
#define TAG_SLOT 0x534c4f54u  /* per-slot placement draws */

static inline double u01(uint32_t r) { return (double)r * 0x1p-32; }
static inline int64_t draw_index(uint32_t r, int64_t n) { return (int64_t)(((uint64_t)r * (uint64_t)n) >> 32); }

/* ------------------------------------------------------------------------ */
/* sin and cos of 2*pi*u                                                     */
/* ------------------------------------------------------------------------ */
static const double SINC[9] = {
    0x1.921fb54442d18p+2, -0x1.4abbce625be53p+5, 0x1.466bc6775aae2p+6,
    -0x1.32d2cce62bd86p+6, 0x1.50783487ee782p+5, -0x1.e3074fde8871fp+3,
    0x1.e8f434d018d63p+1, -0x1.6fadb9f155744p-1, 0x1.aaec32af93359p-4};
static const double COSC[10] = {
    0x1.0000000000000p+0, -0x1.3bd3cc9be45dep+4, 0x1.03c1f081b5ac4p+6,
    -0x1.55d3c7e3cbffap+6, 0x1.e1f506891babbp+5, -0x1.a6d1f2a204a8cp+4,
    0x1.f9d38a3763cc3p+2, -0x1.b6e24f44b128fp+0, 0x1.20c62c2f2d7f5p-2,
    -0x1.2a0c591af8314p-5};

ORC_API void orc_sincos2pi(double u, double *s_out, double *c_out)
{
    double q = rint(4.0 * u);
    double r = (4.0 * u - q) * 0.25;        /* exact, |r| <= 1/8 */
    double z = r * r;
    double ps = SINC[8];
    for (int k = 7; k >= 0; k--) ps = fma(ps, z, SINC[k]);
    double s = r * ps;
    double pc = COSC[9];
    for (int k = 8; k >= 0; k--) pc = fma(pc, z, COSC[k]);
    double c = pc;
    int qm = ((int)q) & 3;
    if (qm == 0) { *s_out = s; *c_out = c; }
    else if (qm == 1) { *s_out = c; *c_out = -s; }
    else if (qm == 2) { *s_out = -s; *c_out = -c; }
    else { *s_out = -c; *c_out = s; }
}

/* eulermat from sines/cosines, src/rigid_body.jl:45-57 (row-major output) */
static void euler_from_sc(double c1, double s1, double c2, double s2, double c3, double s3, double A[9])
{
    A[0] = c2 * c3;               A[1] = -c2 * s3;              A[2] = s2;
    A[3] = c1 * s3 + c3 * s1 * s2; A[4] = c1 * c3 - s1 * s2 * s3; A[5] = -c2 * s1;
    A[6] = s1 * s3 - c1 * c3 * s2; A[7] = c1 * s2 * s3 + c3 * s1; A[8] = c1 * c2;
}

/* eulermat(beta, gamma, theta) in radians with libm, for the reference's closed-form tests */
ORC_API void orc_eulermat(double beta, double gamma, double theta, double A[9])
{
    euler_from_sc(cos(beta), sin(beta), cos(gamma), sin(gamma), cos(theta), sin(theta), A);
}

/* eulermat from angles given in turns (angle = 2*pi*u), as used by the placement */
ORC_API void orc_eulermat_turns(double ub, double ug, double ut, double A[9])
{
    double s1, c1, s2, c2, s3, c3;
    orc_sincos2pi(ub, &s1, &c1);
    orc_sincos2pi(ug, &s2, &c2);
    orc_sincos2pi(ut, &s3, &c3);
    euler_from_sc(c1, s1, c2, s2, c3, s3, A);
}

/* move!: x[i] = A*(x[i]-cm) + newcm, cm = mean(x); src/rigid_body.jl:73-80 */
static void move_with_matrix(double *x, int n, const double newcm[3], const double A[9])
{
    double cx = 0.0, cy = 0.0, cz = 0.0;
    for (int a = 0; a < n; a++) { cx += x[3 * a]; cy += x[3 * a + 1]; cz += x[3 * a + 2]; }
    cx = cx / (double)n; cy = cy / (double)n; cz = cz / (double)n;
    for (int a = 0; a < n; a++) {
        double dx = x[3 * a] - cx, dy = x[3 * a + 1] - cy, dz = x[3 * a + 2] - cz;
        x[3 * a]     = A[0] * dx + A[1] * dy + A[2] * dz + newcm[0];
        x[3 * a + 1] = A[3] * dx + A[4] * dy + A[5] * dz + newcm[1];
        x[3 * a + 2] = A[6] * dx + A[7] * dy + A[8] * dz + newcm[2];
    }
}

ORC_API void orc_move(double *x, int n, const double newcm[3], double beta, double gamma, double theta)
{
    double A[9];
    orc_eulermat(beta, gamma, theta, A);
    move_with_matrix(x, n, newcm, A);
}

/* ------------------------------------------------------------------------ */
/* Unit cell and the minimum-image distance                                  */
/* ------------------------------------------------------------------------ */
typedef struct { double H[9]; double Hi[9]; } orc_cell;   /* column major */

static void cell_setup(const double *cell9, int triclinic, orc_cell *c)
{
    for (int k = 0; k < 9; k++) c->H[k] = cell9[k];
    if (!triclinic) {   /* convert_unitcell(::SVector, m): keep the diagonal, src/Trajectory.jl:80 */
        for (int k = 0; k < 9; k++) if (k != 0 && k != 4 && k != 8) c->H[k] = 0.0;
    }
    const double *H = c->H;
    /* M(r,c) = H[3c+r] */
    double m00 = H[0], m10 = H[1], m20 = H[2], m01 = H[3], m11 = H[4], m21 = H[5], m02 = H[6], m12 = H[7], m22 = H[8];
    double c00 = m11 * m22 - m12 * m21;
    double c01 = m12 * m20 - m10 * m22;
    double c02 = m10 * m21 - m11 * m20;
    double det = m00 * c00 + m01 * c01 + m02 * c02;
    double c10 = m02 * m21 - m01 * m22;
    double c11 = m00 * m22 - m02 * m20;
    double c12 = m01 * m20 - m00 * m21;
    double c20 = m01 * m12 - m02 * m11;
    double c21 = m02 * m10 - m00 * m12;
    double c22 = m00 * m11 - m01 * m10;
    /* inverse(r,c) = cofactor(c,r)/det, stored column major */
    c->Hi[0] = c00 / det; c->Hi[1] = c01 / det; c->Hi[2] = c02 / det;
    c->Hi[3] = c10 / det; c->Hi[4] = c11 / det; c->Hi[5] = c12 / det;
    c->Hi[6] = c20 / det; c->Hi[7] = c21 / det; c->Hi[8] = c22 / det;
}

static inline void minimage(const orc_cell *c, double dx, double dy, double dz, double *ox, double *oy, double *oz)
{
    const double *Hi = c->Hi, *H = c->H;
    double fx = fma(Hi[0], dx, fma(Hi[3], dy, Hi[6] * dz));
    double fy = fma(Hi[1], dx, fma(Hi[4], dy, Hi[7] * dz));
    double fz = fma(Hi[2], dx, fma(Hi[5], dy, Hi[8] * dz));
    double nx = rint(fx), ny = rint(fy), nz = rint(fz);
    *ox = fma(-H[0], nx, fma(-H[3], ny, fma(-H[6], nz, dx)));
    *oy = fma(-H[1], nx, fma(-H[4], ny, fma(-H[7], nz, dy)));
    *oz = fma(-H[2], nx, fma(-H[5], ny, fma(-H[8], nz, dz)));
}

static inline double canon_d2(const orc_cell *c, const double *x, const double *y)
{
    double ox, oy, oz;
    minimage(c, y[0] - x[0], y[1] - x[1], y[2] - x[2], &ox, &oy, &oz);
    return fma(ox, ox, fma(oy, oy, oz * oz));
}

ORC_API double orc_canon_d2(const double *cell9, int triclinic, const double *x, const double *y)
{
    orc_cell c;
    cell_setup(cell9, triclinic, &c);
    return canon_d2(&c, x, y);
}

/* cell_volume: dot(cross(a, b), c), src/mddf.jl:350-352 */
static double cell_volume(const orc_cell *c)
{
    const double *a = c->H, *b = c->H + 3, *cc = c->H + 6;
    double x = a[1] * b[2] - a[2] * b[1];
    double y = a[2] * b[0] - a[0] * b[2];
    double z = a[0] * b[1] - a[1] * b[0];
    return x * cc[0] + y * cc[1] + z * cc[2];
}

/* random_move!: wrap relative to irefatom, rotate about the centre, translate.
 * src/rigid_body.jl:107-137.  x is [n][3], modified in place. */
static void random_move_u(double *x, int n, int iref, const orc_cell *c, const double uc[3], const double ua[3])
{
    const double *xr = x + 3 * iref;
    double rx = xr[0], ry = xr[1], rz = xr[2];
    for (int a = 0; a < n; a++) {
        double ox, oy, oz;
        minimage(c, x[3 * a] - rx, x[3 * a + 1] - ry, x[3 * a + 2] - rz, &ox, &oy, &oz);
        x[3 * a] = rx + ox; x[3 * a + 1] = ry + oy; x[3 * a + 2] = rz + oz;
    }
    const double *H = c->H;
    double newcm[3];
    newcm[0] = H[0] * uc[0] + H[3] * uc[1] + H[6] * uc[2];
    newcm[1] = H[1] * uc[0] + H[4] * uc[1] + H[7] * uc[2];
    newcm[2] = H[2] * uc[0] + H[5] * uc[1] + H[8] * uc[2];
    double A[9];
    orc_eulermat_turns(ua[0], ua[1], ua[2], A);
    move_with_matrix(x, n, newcm, A);
}

ORC_API void orc_random_move(double *x, int n, int iref, const double *cell9, int triclinic,
                             const double uc[3], const double ua[3])
{
    orc_cell c;
    cell_setup(cell9, triclinic, &c);
    random_move_u(x, n, iref, &c, uc, ua);
}

/* ------------------------------------------------------------------------ */
/* Binning helpers                                                           */
/* ------------------------------------------------------------------------ */
ORC_API int64_t orc_setbin(double d, double step)
{
    double b = ceil(d / step);
    int64_t ib = (int64_t)b;
    return ib < 1 ? 1 : ib;
}

/* Julia evaluates x^3 with a literal exponent as x*x*x (Base.literal_pow) */
static inline double cube(double x) { return x * x * x; }

ORC_API double orc_sphericalshellvolume(int64_t i, double step)
{
    double rmin = (double)(i - 1) * step;
    return (4.0 * M_PI / 3.0) * (cube(rmin + step) - cube(rmin));
}

ORC_API double orc_shellradius(int64_t i, double step)
{
    double rmin = (double)(i - 1) * step;
    return pow(0.5 * (cube(rmin + step) + cube(rmin)), 1.0 / 3.0);
}

/* ------------------------------------------------------------------------ */
/* MinimumDistance and update_md                                             */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t within_cutoff;
    int32_t i;        /* molecule-local solute atom */
    int64_t j;        /* global solvent atom */
    double d;
    int32_t ref_atom_within_cutoff;
    int32_t pad;
    double d_ref_atom;
} orc_md;

static inline orc_md md_zero(void)
{
    orc_md m;
    m.within_cutoff = 0; m.i = -1; m.j = -1; m.d = INFINITY;
    m.ref_atom_within_cutoff = 0; m.pad = 0; m.d_ref_atom = INFINITY;
    return m;
}

static inline int md_first_wins(const orc_md *m1, const orc_md *m2)
{
    if (m1->d < m2->d) return 1;
    if (m1->d > m2->d) return 0;
    if (m1->d != m2->d) return 0;   /* NaN guard */
    if (m1->i != m2->i) return m1->i < m2->i;
    return m1->j < m2->j;
}

ORC_API orc_md orc_update_md(orc_md m1, orc_md m2)
{
    orc_md r;
    int ref = m1.ref_atom_within_cutoff || m2.ref_atom_within_cutoff;
    double dref = ref ? fmin(m1.d_ref_atom, m2.d_ref_atom) : INFINITY;
    r = md_first_wins(&m1, &m2) ? m1 : m2;
    r.ref_atom_within_cutoff = ref;
    r.d_ref_atom = dref;
    return r;
}

/* ------------------------------------------------------------------------ */
/* Job description and accumulators                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t nmols_u, nu, nmols_v, nv;
    int32_t autocorrelation;
    int32_t irefatom;              /* 0-based atom of the solvent molecule */
    int32_t nbins;
    int32_t usecutoff;
    double binstep, cutoff, dbulk;
    int32_t n_random_samples;
    int32_t coordination_only;
    int32_t triclinic;
    int32_t use_cells;
    uint64_t seed;
    int32_t ngroups_u, ngroups_v;
    /* CSR maps: solute side over the nu molecule-local atoms, solvent side over
       all nmols_v*nv atoms.  NULL pointers mean default groups (atom types).
       In autocorrelation mode the solvent-side map points into solute groups. */
    const int32_t *gu_ptr, *gu_idx;
    const int32_t *gv_ptr, *gv_idx;
    int32_t nthreads;
    int32_t pad;
} orc_desc;

typedef struct {
    double *md_count, *md_count_random, *rdf_count, *rdf_count_random;   /* [nbins] */
    double *gu, *gv, *gu_random, *gv_random;                              /* [G][nbins] */
    double volume_total;
} orc_raw;

typedef struct {
    double *md_count, *md_count_random, *rdf_count, *rdf_count_random;
    double *gu, *gv, *gu_random, *gv_random;
    double volume_total;
    int64_t ngu, ngv, nbins;
} acc_t;

static int acc_alloc(acc_t *a, const orc_desc *d)
{
    a->nbins = d->nbins; a->ngu = d->ngroups_u; a->ngv = d->ngroups_v;
    a->md_count = calloc(d->nbins, sizeof(double));
    a->md_count_random = calloc(d->nbins, sizeof(double));
    a->rdf_count = calloc(d->nbins, sizeof(double));
    a->rdf_count_random = calloc(d->nbins, sizeof(double));
    a->gu = calloc((size_t)d->ngroups_u * d->nbins, sizeof(double));
    a->gv = calloc((size_t)d->ngroups_v * d->nbins, sizeof(double));
    a->gu_random = calloc((size_t)d->ngroups_u * d->nbins, sizeof(double));
    a->gv_random = calloc((size_t)d->ngroups_v * d->nbins, sizeof(double));
    a->volume_total = 0.0;
    return (a->md_count && a->md_count_random && a->rdf_count && a->rdf_count_random &&
            a->gu && a->gv && a->gu_random && a->gv_random) ? 0 : -1;
}

static void acc_free(acc_t *a)
{
    free(a->md_count); free(a->md_count_random); free(a->rdf_count); free(a->rdf_count_random);
    free(a->gu); free(a->gv); free(a->gu_random); free(a->gv_random);
}

/* sum!, src/results.jl:629-649 */
static void acc_sum(acc_t *r1, const acc_t *r2)
{
    int64_t nb = r1->nbins;
    for (int64_t b = 0; b < nb; b++) {
        r1->md_count[b] += r2->md_count[b];
        r1->md_count_random[b] += r2->md_count_random[b];
    }
    for (int64_t k = 0; k < r1->ngu * nb; k++) { r1->gu[k] += r2->gu[k]; r1->gu_random[k] += r2->gu_random[k]; }
    for (int64_t k = 0; k < r1->ngv * nb; k++) { r1->gv[k] += r2->gv[k]; r1->gv_random[k] += r2->gv_random[k]; }
    for (int64_t b = 0; b < nb; b++) {
        r1->rdf_count[b] += r2->rdf_count[b];
        r1->rdf_count_random[b] += r2->rdf_count_random[b];
    }
    r1->volume_total += r2->volume_total;
}

/* update_group_count!, src/update_counters.jl:22-36 */
static inline void group_add(double *gc, int64_t nbins, int64_t ibin, int64_t atom, double w,
                             const int32_t *ptr, const int32_t *idx, int64_t natomspermol)
{
    if (!ptr) {
        int64_t t = atom % natomspermol;
        gc[t * nbins + ibin] += w;
    } else {
        for (int32_t k = ptr[atom]; k < ptr[atom + 1]; k++) gc[(int64_t)idx[k] * nbins + ibin] += w;
    }
}

/* update_counters!, src/update_counters.jl:43-88 */
static void update_counters(const orc_desc *d, acc_t *a, const orc_md *list, double w, int random)
{
    int64_t nb = d->nbins;
    double *md = random ? a->md_count_random : a->md_count;
    double *rdf = random ? a->rdf_count_random : a->rdf_count;
    double *gu = random ? a->gu_random : a->gu;
    double *gv = random ? a->gv_random : a->gv;
    for (int64_t v = 0; v < d->nmols_v; v++) {
        const orc_md *m = &list[v];
        if (!m->within_cutoff) continue;
        int64_t ibin = orc_setbin(m->d, d->binstep) - 1;
        md[ibin] += w;
        if (d->autocorrelation) {
            group_add(gu, nb, ibin, m->i, w / 2, d->gu_ptr, d->gu_idx, d->nu);
            group_add(gu, nb, ibin, m->j, w / 2, d->gv_ptr, d->gv_idx, d->nu);
        } else {
            group_add(gu, nb, ibin, m->i, w, d->gu_ptr, d->gu_idx, d->nu);
            group_add(gv, nb, ibin, m->j, w, d->gv_ptr, d->gv_idx, d->nv);
        }
        if (m->ref_atom_within_cutoff) {
            int64_t rb = orc_setbin(m->d_ref_atom, d->binstep) - 1;
            rdf[rb] += w;
        }
    }
}

static inline int inbulk(const orc_desc *d, const orc_md *m)
{
    return d->usecutoff ? (m->within_cutoff && m->d > d->dbulk) : !m->within_cutoff;
}

/* ------------------------------------------------------------------------ */
/* Solvent cell list (CellListMap stand-in: enumerates candidates only; the  */
/* distance that is recorded is always the canonical one)                    */
/* ------------------------------------------------------------------------ */
typedef struct {
    int nc[3];
    int *start;     /* [ncell+1] */
    int *atoms;     /* [natoms] sorted by cell */
    double *xw;     /* [natoms][3] wrapped coordinates (indexed by atom) */
    int natoms, cap_cells, cap_atoms;
} celllist;

static void cl_free(celllist *cl) { free(cl->start); free(cl->atoms); free(cl->xw); memset(cl, 0, sizeof(*cl)); }

static void plane_spacings(const orc_cell *c, double sp[3])
{
    double V = fabs(cell_volume(c));
    for (int k = 0; k < 3; k++) {
        const double *p = c->H + 3 * ((k + 1) % 3), *q = c->H + 3 * ((k + 2) % 3);
        double x = p[1] * q[2] - p[2] * q[1], y = p[2] * q[0] - p[0] * q[2], z = p[0] * q[1] - p[1] * q[0];
        sp[k] = V / sqrt(x * x + y * y + z * z);
    }
}

static inline void wrap_frac(const orc_cell *c, const double *x, double f[3])
{
    const double *Hi = c->Hi;
    f[0] = Hi[0] * x[0] + Hi[3] * x[1] + Hi[6] * x[2];
    f[1] = Hi[1] * x[0] + Hi[4] * x[1] + Hi[7] * x[2];
    f[2] = Hi[2] * x[0] + Hi[5] * x[1] + Hi[8] * x[2];
    for (int k = 0; k < 3; k++) { f[k] -= floor(f[k]); if (f[k] >= 1.0) f[k] = 0.0; }
}

static int cl_build(celllist *cl, const orc_cell *c, const double *y, int natoms, double side)
{
    double sp[3];
    plane_spacings(c, sp);
    for (int k = 0; k < 3; k++) { int n = (int)floor(sp[k] / side); cl->nc[k] = n < 1 ? 1 : n; }
    int ncell = cl->nc[0] * cl->nc[1] * cl->nc[2];
    if (ncell + 1 > cl->cap_cells) { free(cl->start); cl->start = malloc(sizeof(int) * (ncell + 1)); cl->cap_cells = ncell + 1; }
    if (natoms > cl->cap_atoms) {
        free(cl->atoms); free(cl->xw);
        cl->atoms = malloc(sizeof(int) * natoms); cl->xw = malloc(sizeof(double) * 3 * natoms);
        cl->cap_atoms = natoms;
    }
    if (!cl->start || !cl->atoms || !cl->xw) return -1;
    cl->natoms = natoms;
    int *cellof = malloc(sizeof(int) * (natoms > 0 ? natoms : 1));
    if (!cellof) return -1;
    memset(cl->start, 0, sizeof(int) * (ncell + 1));
    const double *H = c->H;
    for (int j = 0; j < natoms; j++) {
        double f[3];
        wrap_frac(c, y + 3 * j, f);
        int ci[3];
        for (int k = 0; k < 3; k++) { ci[k] = (int)(f[k] * cl->nc[k]); if (ci[k] >= cl->nc[k]) ci[k] = cl->nc[k] - 1; }
        int cell = (ci[2] * cl->nc[1] + ci[1]) * cl->nc[0] + ci[0];
        cellof[j] = cell;
        cl->start[cell + 1]++;
        cl->xw[3 * j]     = H[0] * f[0] + H[3] * f[1] + H[6] * f[2];
        cl->xw[3 * j + 1] = H[1] * f[0] + H[4] * f[1] + H[7] * f[2];
        cl->xw[3 * j + 2] = H[2] * f[0] + H[5] * f[1] + H[8] * f[2];
    }
    for (int k = 0; k < ncell; k++) cl->start[k + 1] += cl->start[k];
    int *fill = calloc(ncell > 0 ? ncell : 1, sizeof(int));
    if (!fill) { free(cellof); return -1; }
    for (int j = 0; j < natoms; j++) { int cell = cellof[j]; cl->atoms[cl->start[cell] + fill[cell]++] = j; }
    free(fill);
    free(cellof);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* minimum_distances!, src/minimum_distances.jl:81-148                       */
/* ------------------------------------------------------------------------ */
static inline void offer(orc_md *list, const orc_desc *d, int i, int64_t j, double d2, double cut2)
{
    if (!(d2 <= cut2)) return;
    double dist = sqrt(d2);
    int ref = (int)(j % d->nv) == d->irefatom;
    orc_md cand;
    cand.within_cutoff = 1; cand.i = i; cand.j = j; cand.d = dist;
    cand.ref_atom_within_cutoff = ref; cand.pad = 0; cand.d_ref_atom = ref ? dist : INFINITY;
    int64_t v = j / d->nv;
    list[v] = orc_update_md(list[v], cand);
}

static void minimum_distances(const orc_desc *d, const orc_cell *c, const double *xmol /* [nu][3] */,
                              const double *y /* [Nv][3] */, int isolute, orc_md *list,
                              const celllist *cl, double cut)
{
    int64_t Nv = (int64_t)d->nmols_v * d->nv;
    double cut2 = cut * cut;
    for (int64_t v = 0; v < d->nmols_v; v++) list[v] = md_zero();
    int self = d->autocorrelation;
    if (!cl) {
        for (int i = 0; i < d->nu; i++) {
            const double *x = xmol + 3 * i;
            for (int64_t j = 0; j < Nv; j++) {
                if (self && j / d->nv == isolute) continue;
                offer(list, d, i, j, canon_d2(c, x, y + 3 * j), cut2);
            }
        }
        return;
    }
    const double *H = c->H;
    double filt = cut2 * (1.0 + 1e-9) + 1e-9;
    for (int i = 0; i < d->nu; i++) {
        double f[3], xw[3];
        wrap_frac(c, xmol + 3 * i, f);
        int ci[3];
        for (int k = 0; k < 3; k++) { ci[k] = (int)(f[k] * cl->nc[k]); if (ci[k] >= cl->nc[k]) ci[k] = cl->nc[k] - 1; }
        xw[0] = H[0] * f[0] + H[3] * f[1] + H[6] * f[2];
        xw[1] = H[1] * f[0] + H[4] * f[1] + H[7] * f[2];
        xw[2] = H[2] * f[0] + H[5] * f[1] + H[8] * f[2];
        for (int oz = -1; oz <= 1; oz++)
        for (int oy = -1; oy <= 1; oy++)
        for (int ox = -1; ox <= 1; ox++) {
            int o[3] = {ox, oy, oz}, cc[3];
            double m[3];
            for (int k = 0; k < 3; k++) {
                int t = ci[k] + o[k];
                int mk = t < 0 ? -1 : (t >= cl->nc[k] ? 1 : 0);
                cc[k] = t - mk * cl->nc[k];
                m[k] = (double)mk;
            }
            double sx = H[0] * m[0] + H[3] * m[1] + H[6] * m[2];
            double sy = H[1] * m[0] + H[4] * m[1] + H[7] * m[2];
            double sz = H[2] * m[0] + H[5] * m[1] + H[8] * m[2];
            int cell = (cc[2] * cl->nc[1] + cc[1]) * cl->nc[0] + cc[0];
            for (int p = cl->start[cell]; p < cl->start[cell + 1]; p++) {
                int64_t j = cl->atoms[p];
                if (self && j / d->nv == isolute) continue;
                double dx = cl->xw[3 * j] + sx - xw[0];
                double dy = cl->xw[3 * j + 1] + sy - xw[1];
                double dz = cl->xw[3 * j + 2] + sz - xw[2];
                double fd2 = dx * dx + dy * dy + dz * dz;
                if (fd2 > filt) continue;
                offer(list, d, i, j, canon_d2(c, xmol + 3 * i, y + 3 * j), cut2);
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* One frame: mddf_frame! / coordination_number_frame!                       */
/* ------------------------------------------------------------------------ */
typedef struct {
    orc_md *list, *list_r;
    int64_t *bulk;
    double *placed;        /* [Nv][3] */
    celllist cl, clr;
} work_t;

static int work_alloc(work_t *w, const orc_desc *d)
{
    memset(w, 0, sizeof(*w));
    int64_t Nv = (int64_t)d->nmols_v * d->nv;
    w->list = malloc(sizeof(orc_md) * (d->nmols_v > 0 ? d->nmols_v : 1));
    w->list_r = malloc(sizeof(orc_md) * (d->nmols_v > 0 ? d->nmols_v : 1));
    w->bulk = malloc(sizeof(int64_t) * (d->nmols_v > 0 ? d->nmols_v : 1));
    w->placed = malloc(sizeof(double) * 3 * (Nv > 0 ? Nv : 1));
    return (w->list && w->list_r && w->bulk && w->placed) ? 0 : -1;
}

static void work_free(work_t *w)
{
    free(w->list); free(w->list_r); free(w->bulk); free(w->placed);
    cl_free(&w->cl); cl_free(&w->clr);
}

/* randomize_solvent!: src/mddf.jl:65-88 */
static void randomize_solvent(const orc_desc *d, const orc_cell *c, const double *yread, int64_t nbulk,
                              const int64_t *bulk, uint32_t frame, uint32_t sample, double *placed)
{
    uint32_t key[2] = {(uint32_t)d->seed, (uint32_t)(d->seed >> 32)};
    int nv = d->nv;
    for (int64_t v = 0; v < d->nmols_v; v++) {
        uint32_t ctr[4] = {(uint32_t)v, sample, frame, TAG_SLOT}, r[4];
        orc_philox4x32_10(ctr, key, r);
        uint32_t ctr2[4] = {(uint32_t)v, sample, frame, TAG_SLOT + 1u}, r2[4];
        orc_philox4x32_10(ctr2, key, r2);
        int64_t jmol = nbulk > 0 ? bulk[draw_index(r[0], nbulk)] : draw_index(r[0], d->nmols_v);
        double *x = placed + 3 * v * nv;
        memcpy(x, yread + 3 * jmol * nv, sizeof(double) * 3 * nv);
        double uc[3] = {u01(r[1]), u01(r[2]), u01(r[3])};
        double ua[3] = {u01(r2[0]), u01(r2[1]), u01(r2[2])};
        random_move_u(x, nv, d->irefatom, c, uc, ua);
    }
}

static int do_frame(const orc_desc *d, acc_t *a, work_t *wk, const double *xu, const double *yv,
                    const double *cell9, double w, uint32_t frame)
{
    orc_cell c;
    cell_setup(cell9, d->triclinic, &c);
    a->volume_total += w * cell_volume(&c);
    double cut = d->usecutoff ? d->cutoff : d->dbulk;
    int64_t Nv = (int64_t)d->nmols_v * d->nv;
    int nrand = d->coordination_only ? 0 : d->n_random_samples;
    uint32_t key[2] = {(uint32_t)d->seed, (uint32_t)(d->seed >> 32)};
    int64_t refu[4096];
    int64_t *ref_solutes = nrand <= 4096 ? refu : malloc(sizeof(int64_t) * nrand);
    if (!ref_solutes) return -1;
    for (int k = 0; k < nrand; k++) {
        uint32_t ctr[4] = {(uint32_t)k, 0u, frame, TAG_REF}, r[4];
        orc_philox4x32_10(ctr, key, r);
        ref_solutes[k] = draw_index(r[0], d->nmols_u);
    }
    const celllist *clp = NULL;
    if (d->use_cells) {
        if (cl_build(&wk->cl, &c, yv, (int)Nv, cut) != 0) return -1;
        clp = &wk->cl;
    }
    for (int isolute = 0; isolute < d->nmols_u; isolute++) {
        const double *xmol = xu + 3 * (int64_t)isolute * d->nu;
        minimum_distances(d, &c, xmol, yv, isolute, wk->list, clp, cut);
        update_counters(d, a, wk->list, w, 0);
        int any = 0;
        for (int k = 0; k < nrand; k++) if (ref_solutes[k] == isolute) { any = 1; break; }
        if (!any) continue;
        int64_t nbulk = 0;
        for (int64_t v = 0; v < d->nmols_v; v++) {
            if (d->autocorrelation && v == isolute) continue;
            if (inbulk(d, &wk->list[v])) wk->bulk[nbulk++] = v;
        }
        for (int k = 0; k < nrand; k++) {
            if (ref_solutes[k] != isolute) continue;
            randomize_solvent(d, &c, yv, nbulk, wk->bulk, frame, (uint32_t)k, wk->placed);
            const celllist *clr = NULL;
            if (d->use_cells) {
                if (cl_build(&wk->clr, &c, wk->placed, (int)Nv, cut) != 0) return -1;
                clr = &wk->clr;
            }
            minimum_distances(d, &c, xmol, wk->placed, isolute, wk->list_r, clr, cut);
            update_counters(d, a, wk->list_r, w, 1);
        }
    }
    if (ref_solutes != refu) free(ref_solutes);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Job driver with the reference's chunked parallel setup                    */
/* ------------------------------------------------------------------------ */
typedef struct {
    const orc_desc *d;
    int nframes;
    const int64_t *frame_index;
    const double *xu, *xv, *cells, *w;
    int f0, f1;
    acc_t acc;
    int status;
} chunk_t;

static void *chunk_run(void *arg)
{
    chunk_t *ch = (chunk_t *)arg;
    const orc_desc *d = ch->d;
    work_t wk;
    ch->status = 0;
    if (work_alloc(&wk, d) != 0) { ch->status = -1; work_free(&wk); return NULL; }
    int64_t Nu = (int64_t)d->nmols_u * d->nu, Nv = (int64_t)d->nmols_v * d->nv;
    for (int f = ch->f0; f < ch->f1; f++) {
        if (do_frame(d, &ch->acc, &wk, ch->xu + 3 * Nu * f, ch->xv + 3 * Nv * f, ch->cells + 9 * f,
                     ch->w[f], (uint32_t)ch->frame_index[f]) != 0) { ch->status = -1; break; }
    }
    work_free(&wk);
    return NULL;
}

/* Runs the frames given (already selected by firstframe/stride/weights) and
 * returns raw accumulators.  frame_index holds the 1-based trajectory frame
 * numbers, which key the random streams. */
ORC_API int orc_run(const orc_desc *d, int nframes, const int64_t *frame_index, const double *xu,
                    const double *xv, const double *cells, const double *w, orc_raw *out)
{
    int nth = d->nthreads > 0 ? d->nthreads : 1;
    if (nth > nframes) nth = nframes > 0 ? nframes : 1;
    chunk_t *ch = calloc(nth, sizeof(chunk_t));
    pthread_t *th = calloc(nth, sizeof(pthread_t));
    if (!ch || !th) { free(ch); free(th); return -1; }
    int status = 0;
    for (int t = 0; t < nth; t++) {
        ch[t].d = d; ch[t].nframes = nframes; ch[t].frame_index = frame_index;
        ch[t].xu = xu; ch[t].xv = xv; ch[t].cells = cells; ch[t].w = w;
        /* ChunkSplitters-style contiguous chunks */
        ch[t].f0 = (int)((int64_t)nframes * t / nth);
        ch[t].f1 = (int)((int64_t)nframes * (t + 1) / nth);
        if (acc_alloc(&ch[t].acc, d) != 0) status = -1;
    }
    if (status == 0) {
        if (nth == 1) chunk_run(&ch[0]);
        else {
            for (int t = 0; t < nth; t++) pthread_create(&th[t], NULL, chunk_run, &ch[t]);
            for (int t = 0; t < nth; t++) pthread_join(th[t], NULL);
        }
        acc_t total;
        if (acc_alloc(&total, d) != 0) status = -1;
        else {
            for (int t = 0; t < nth; t++) { if (ch[t].status != 0) status = -1; acc_sum(&total, &ch[t].acc); }
            int64_t nb = d->nbins;
            memcpy(out->md_count, total.md_count, sizeof(double) * nb);
            memcpy(out->md_count_random, total.md_count_random, sizeof(double) * nb);
            memcpy(out->rdf_count, total.rdf_count, sizeof(double) * nb);
            memcpy(out->rdf_count_random, total.rdf_count_random, sizeof(double) * nb);
            memcpy(out->gu, total.gu, sizeof(double) * nb * d->ngroups_u);
            memcpy(out->gv, total.gv, sizeof(double) * nb * d->ngroups_v);
            memcpy(out->gu_random, total.gu_random, sizeof(double) * nb * d->ngroups_u);
            memcpy(out->gv_random, total.gv_random, sizeof(double) * nb * d->ngroups_v);
            out->volume_total = total.volume_total;
            acc_free(&total);
        }
    }
    for (int t = 0; t < nth; t++) acc_free(&ch[t].acc);
    free(ch); free(th);
    return status;
}

/* Per-molecule records of one real pass (for record-level parity checks). */
ORC_API int orc_records(const orc_desc *d, const double *xu, const double *xv, const double *cell9,
                        int isolute, orc_md *list)
{
    orc_cell c;
    cell_setup(cell9, d->triclinic, &c);
    double cut = d->usecutoff ? d->cutoff : d->dbulk;
    celllist cl;
    memset(&cl, 0, sizeof(cl));
    const celllist *clp = NULL;
    if (d->use_cells) {
        if (cl_build(&cl, &c, xv, d->nmols_v * d->nv, cut) != 0) { cl_free(&cl); return -1; }
        clp = &cl;
    }
    minimum_distances(d, &c, xu + 3 * (int64_t)isolute * d->nu, xv, isolute, list, clp, cut);
    cl_free(&cl);
    return 0;
}

/* Placed solvent coordinates of one random pass (frame, sample) given the
 * bulk list, for checks of the placement itself. */
ORC_API void orc_placement(const orc_desc *d, const double *xv, const double *cell9, int64_t nbulk,
                           const int64_t *bulk, uint32_t frame, uint32_t sample, double *placed)
{
    orc_cell c;
    cell_setup(cell9, d->triclinic, &c);
    randomize_solvent(d, &c, xv, nbulk, bulk, frame, sample, placed);
}

ORC_API int64_t orc_ref_solute(uint64_t seed, uint32_t frame, uint32_t sample, int64_t nmols_u)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {sample, 0u, frame, TAG_REF}, r[4];
    orc_philox4x32_10(ctr, key, r);
    return draw_index(r[0], nmols_u);
}

/* ------------------------------------------------------------------------ */
/* finalresults!, src/results.jl:311-469                                     */
/* ------------------------------------------------------------------------ */
typedef struct {
    double *d, *md_count, *md_count_random, *coordination_number, *coordination_number_random;
    double *mddf, *kb;
    double *gu, *gv, *gu_random, *gv_random;
    double *rdf_count, *rdf_count_random, *sum_rdf_count, *sum_rdf_count_random, *rdf, *kb_rdf;
    double density_solute, density_solvent, density_solvent_bulk;
    double volume_total, volume_bulk, volume_domain;
    double *volume_shell;
} orc_final;

/* raw and fin may alias arrays; fin arrays must be zero-initialised. */
ORC_API void orc_finalize(const orc_desc *d, const orc_raw *raw, double Q, orc_final *fin)
{
    const double Angs3tocm3permol = 6.022140857e23 / 1e24;
    int64_t nb = d->nbins;
    double step = d->binstep;
    for (int64_t b = 0; b < nb; b++) fin->d[b] = orc_shellradius(b + 1, step);
    double nu = (double)d->nmols_u;
    int64_t ngu = d->ngroups_u, ngv = d->ngroups_v;
    for (int64_t b = 0; b < nb; b++) fin->md_count[b] = raw->md_count[b] / (nu * Q);
    for (int64_t k = 0; k < ngu * nb; k++) fin->gu[k] = raw->gu[k] / (nu * Q);
    if (d->coordination_only) {
        if (d->autocorrelation) {
            for (int64_t k = 0; k < ngv * nb && k < ngu * nb; k++) fin->gv[k] = fin->gu[k];
        } else {
            for (int64_t k = 0; k < ngv * nb; k++) fin->gv[k] = raw->gv[k] / (nu * Q);
        }
        for (int64_t b = 0; b < nb; b++) fin->rdf_count[b] = raw->rdf_count[b] / (nu * Q);
        fin->volume_total = raw->volume_total / Q;
        fin->density_solvent = (double)d->nmols_v / fin->volume_total;
        fin->density_solute = nu / fin->volume_total;
        double s = 0.0;
        for (int64_t b = 0; b < nb; b++) { s += fin->md_count[b]; fin->coordination_number[b] = s; }
        s = 0.0;
        for (int64_t b = 0; b < nb; b++) { s += fin->rdf_count[b]; fin->sum_rdf_count[b] = s; }
        return;
    }
    double srand_ = (double)d->n_random_samples;
    double solvent_nmols = d->autocorrelation ? (double)(d->nmols_v - 1) : (double)d->nmols_v;
    for (int64_t k = 0; k < ngu * nb; k++) fin->gu_random[k] = raw->gu_random[k] / (srand_ * Q);
    if (d->autocorrelation) {
        for (int64_t k = 0; k < ngv * nb && k < ngu * nb; k++) { fin->gv[k] = fin->gu[k]; fin->gv_random[k] = fin->gu_random[k]; }
    } else {
        for (int64_t k = 0; k < ngv * nb; k++) {
            fin->gv[k] = raw->gv[k] / (nu * Q);
            fin->gv_random[k] = raw->gv_random[k] / (srand_ * Q);
        }
    }
    for (int64_t b = 0; b < nb; b++) {
        fin->md_count_random[b] = raw->md_count_random[b] / (srand_ * Q);
        fin->rdf_count[b] = raw->rdf_count[b] / (nu * Q);
        fin->rdf_count_random[b] = raw->rdf_count_random[b] / (srand_ * Q);
    }
    fin->volume_total = raw->volume_total / Q;
    for (int64_t b = 0; b < nb; b++) fin->volume_shell[b] = fin->volume_total * (fin->rdf_count_random[b] / solvent_nmols);
    int64_t ibulk = orc_setbin(d->dbulk + 0.5 * step, step);
    double dom = 0.0;
    for (int64_t b = 0; b < ibulk - 1; b++) dom += fin->volume_shell[b];
    fin->volume_domain = dom;
    double n_solvent_in_bulk;
    if (!d->usecutoff) {
        fin->volume_bulk = fin->volume_total - fin->volume_domain;
        double s = 0.0;
        for (int64_t b = 0; b < nb; b++) s += fin->rdf_count[b];
        n_solvent_in_bulk = solvent_nmols - s;
    } else {
        double s = 0.0, vb = 0.0;
        for (int64_t b = ibulk - 1; b < nb; b++) { s += fin->rdf_count[b]; vb += fin->volume_shell[b]; }
        n_solvent_in_bulk = s;
        fin->volume_bulk = vb;
    }
    fin->density_solvent = (double)d->nmols_v / fin->volume_total;
    fin->density_solute = nu / fin->volume_total;
    fin->density_solvent_bulk = n_solvent_in_bulk / fin->volume_bulk;
    double density_fix = fin->density_solvent_bulk / fin->density_solvent;
    /* renormalize!, src/results.jl:378-428 */
    for (int64_t b = 0; b < nb; b++) { fin->md_count_random[b] *= density_fix; fin->rdf_count_random[b] *= density_fix; }
    for (int64_t k = 0; k < ngu * nb; k++) fin->gu_random[k] *= density_fix;
    for (int64_t k = 0; k < ngv * nb; k++) fin->gv_random[k] *= density_fix;
    double s1 = 0.0, s2 = 0.0;
    for (int64_t b = 0; b < nb; b++) {
        s1 += fin->md_count[b]; fin->coordination_number[b] = s1;
        s2 += fin->md_count_random[b]; fin->coordination_number_random[b] = s2;
    }
    for (int64_t b = 0; b < nb; b++) {
        if (fin->md_count_random[b] > 0.0) fin->mddf[b] = fin->md_count[b] / fin->md_count_random[b];
        fin->kb[b] = Angs3tocm3permol * (1.0 / fin->density_solvent_bulk) *
                     (fin->coordination_number[b] - fin->coordination_number_random[b]);
        if (fin->rdf_count_random[b] > 0.0) fin->rdf[b] = fin->rdf_count[b] / fin->rdf_count_random[b];
        if (b == 0) {
            fin->sum_rdf_count[b] = fin->rdf_count[b];
            fin->sum_rdf_count_random[b] = fin->rdf_count_random[b];
        } else {
            fin->sum_rdf_count[b] = fin->sum_rdf_count[b - 1] + fin->rdf_count[b];
            fin->sum_rdf_count_random[b] = fin->sum_rdf_count_random[b - 1] + fin->rdf_count_random[b];
        }
        fin->kb_rdf[b] = Angs3tocm3permol * (1.0 / fin->density_solvent_bulk) *
                         (fin->sum_rdf_count[b] - fin->sum_rdf_count_random[b]);
    }
}
