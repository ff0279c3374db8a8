/*
 * imexlb_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C, FP64, operation-by-operation CPU restatement of the IMEXLB-1.0
 * lattice Boltzmann time step at one MPI rank.  It exists so that the HIP
 * product path can be checked against the reference's algorithm.  Only
 * tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may load it,
 * and only as the checker; the product library never links or calls it.
 *
 * Layout follows the reference exactly: one flat array per field with a
 * ghost layer of width 3 on every side (cart_mpi.f90:10,133-136), node id
 * (x+3) + dy*(y+3) [+ dz*(z+3)], populations stored AoS as f[id*nq+iq]
 * (lbm.f90:44), velocity AoS as u[id*dim+d].  Every loop keeps the
 * reference's iteration order and every expression keeps its evaluation
 * order, so that a build without floating-point contraction
 * (-ffp-contract=off, no fast-math) reproduces its IEEE-754 results.
 *
 * Where the reference reads outside its arrays (undefined behaviour there),
 * this restatement reads 0 and counts the event in `ub_reads`, so a caller
 * can tell whether a result depends on it.
 *
 * Parity pinning (what fixes each function independently of this text) is
 * listed beside each function and enforced by tests/test_oracle_*.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_GHOST 3

typedef struct orc {
    int dim, nq;
    int n[3];          /* nx, ny, nz: largest node index per axis            */
    int L[3];          /* local_length = n+1 at one rank                      */
    int64_t A;         /* array_size = prod(L+6)                              */
    int64_t dy, dz;    /* flat strides of one y / z step                      */
    int e[27 * 3];     /* lattice velocities, e[iq*dim+d]                     */
    double t[27];      /* lattice weights                                     */
    double ome, rho0, uu, pos, radius;
    double *f, *fb, *p, *u;
    int32_t *geo;
    int64_t *fluid_id, *bl, *br, *bu, *bd, *bf, *bb, *b_user;
    int64_t size_fluid, l_size, r_size, u_size, d_size, f_size, b_size, user_size;
    int64_t ub_reads;  /* out-of-array reads (reference UB) that were needed  */
} orc;

/* ------------------------------------------------------------------------
 * a1  DataInput derived quantities   (initialization.f90:59-81;
 *                                     initialization_3d.f90:60-82)
 * out: [0]charlength [1]pos [2]nu [3]t_intv [4]uu [5]tau [6]ma [7]radius
 * iout: [0]interv [1]max_step  (Fortran real->integer assignment truncates)
 * Pinned by: the derived-value table computed in IEEE double for five inputs
 * (tests/test_oracle_setup.py::test_data_input_known_values).
 * ---------------------------------------------------------------------- */
void orc_data_input(int nx, double Lc, double Xpos, double Re, double ome,
                    double dt, double rho0, double max_t, double inter,
                    double out[8], int64_t iout[2])
{
    (void)rho0;
    double charlength = Lc * (double)nx;
    double pos = Xpos * (double)nx;
    double nu = (2.0 - ome) * dt / (6.0 * ome);
    double t_intv = charlength * charlength / nu;
    int32_t interv = (int32_t)(inter * t_intv);
    int32_t max_step = (int32_t)(max_t * t_intv);
    double uu = Re * nu / charlength;
    double tau = 3.0 * nu;
    double ma = uu * sqrt(3.0);
    double radius = charlength * 0.5;
    out[0] = charlength; out[1] = pos; out[2] = nu; out[3] = t_intv;
    out[4] = uu; out[5] = tau; out[6] = ma; out[7] = radius;
    iout[0] = interv; iout[1] = max_step;
}

/* ------------------------------------------------------------------------
 * a2  InitLattice   (initialization.f90:112-123; initialization_3d.f90:108-122)
 * 2D: iq=(i+1)*3+(j+1); t = 4*(1-0.75 i^2)*(1-0.75 j^2)/9 with the product
 * formed in single precision (exact for i,j in {-1,0,1}) then divided in
 * double.  3D: iq=(k+1)*9+(i+1)*3+(j+1); t = 8/4^(|i|+|j|+|k|)/27.
 * Pinned by: the direction picture at initialization.f90:106-111 and the
 * isotropy identities sum t = 1, sum t e = 0, sum t e e = delta/3,
 * sum t e^4 = 1/3 (tests/test_oracle_setup.py).
 * ---------------------------------------------------------------------- */
void orc_init_lattice(int dim, int *e, double *t)
{
    if (dim == 2) {
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i) {
                int iq = (i + 1) * 3 + (j + 1);
                e[iq * 2] = i;
                e[iq * 2 + 1] = j;
                float w = 4.0f * (1.0f - 0.75f * (float)(i * i)) * (1.0f - 0.75f * (float)(j * j));
                t[iq] = (double)w / 9.0;
            }
    } else {
        for (int j = -1; j <= 1; ++j)
            for (int i = -1; i <= 1; ++i)
                for (int k = -1; k <= 1; ++k) {
                    int iq = (k + 1) * 9 + (i + 1) * 3 + (j + 1);
                    e[iq * 3] = i;
                    e[iq * 3 + 1] = j;
                    e[iq * 3 + 2] = k;
                    double p4 = 1.0;
                    for (int m = 0; m < abs(i) + abs(j) + abs(k); ++m) p4 *= 4.0;
                    t[iq] = 8.0 / p4 / 27.0;
                }
    }
}

/* ------------------------------------------------------------------------
 * a3  SetCartesian + SetLocalLength + AllocateArrays at one rank
 *     (cart_mpi.f90:57-138, initialization.f90:127-144; 3D counterparts)
 * local_length = n+1, local_start = 0, array_size = prod(L+6).  Arrays are
 * zero-filled: the reference never writes f/fb at solid nodes and reads the
 * zero pages the allocator hands out there.
 * ---------------------------------------------------------------------- */
orc *orc_create(int dim, int nx, int ny, int nz)
{
    if (dim != 2 && dim != 3) return NULL;
    orc *o = (orc *)calloc(1, sizeof(orc));
    if (!o) return NULL;
    o->dim = dim;
    o->nq = dim == 2 ? 9 : 27;
    o->n[0] = nx; o->n[1] = ny; o->n[2] = dim == 3 ? nz : 0;
    for (int d = 0; d < 3; ++d) o->L[d] = o->n[d] + 1;
    o->A = 1;
    for (int d = 0; d < dim; ++d) o->A *= (int64_t)(o->L[d] + 2 * ORC_GHOST);
    o->dy = o->L[0] + 2 * ORC_GHOST;
    o->dz = dim == 3 ? o->dy * (o->L[1] + 2 * ORC_GHOST) : 0;
    orc_init_lattice(dim, o->e, o->t);
    int64_t A = o->A, nq = o->nq;
    o->f = (double *)calloc((size_t)(A * nq), sizeof(double));
    o->fb = (double *)calloc((size_t)(A * nq), sizeof(double));
    o->p = (double *)calloc((size_t)A, sizeof(double));
    o->u = (double *)calloc((size_t)(A * dim), sizeof(double));
    o->geo = (int32_t *)calloc((size_t)A, sizeof(int32_t));
    o->fluid_id = (int64_t *)calloc((size_t)A, sizeof(int64_t));
    o->b_user = (int64_t *)calloc((size_t)A, sizeof(int64_t));
    int64_t face = dim == 2 ? (int64_t)(o->L[1] + 3) : (int64_t)(o->L[1] + 6) * (o->L[2] + 6);
    int64_t face_x = dim == 2 ? (int64_t)(o->L[0] + 3) : (int64_t)(o->L[0] + 6) * (o->L[1] + 6);
    int64_t face_y = dim == 2 ? 1 : (int64_t)(o->L[0] + 6) * (o->L[2] + 6);
    o->bl = (int64_t *)calloc((size_t)face, sizeof(int64_t));
    o->br = (int64_t *)calloc((size_t)face, sizeof(int64_t));
    o->bu = (int64_t *)calloc((size_t)face_x, sizeof(int64_t));
    o->bd = (int64_t *)calloc((size_t)face_x, sizeof(int64_t));
    o->bf = (int64_t *)calloc((size_t)face_y, sizeof(int64_t));
    o->bb = (int64_t *)calloc((size_t)face_y, sizeof(int64_t));
    if (!o->f || !o->fb || !o->p || !o->u || !o->geo || !o->fluid_id || !o->b_user ||
        !o->bl || !o->br || !o->bu || !o->bd || !o->bf || !o->bb) {
        free(o->f); free(o->fb); free(o->p); free(o->u); free(o->geo);
        free(o->fluid_id); free(o->b_user); free(o->bl); free(o->br);
        free(o->bu); free(o->bd); free(o->bf); free(o->bb); free(o);
        return NULL;
    }
    return o;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    free(o->f); free(o->fb); free(o->p); free(o->u); free(o->geo);
    free(o->fluid_id); free(o->b_user); free(o->bl); free(o->br);
    free(o->bu); free(o->bd); free(o->bf); free(o->bb);
    free(o);
}

void orc_set_params(orc *o, double ome, double rho0, double uu)
{
    o->ome = ome; o->rho0 = rho0; o->uu = uu;
}

/* flat id of the node with coordinates (x,y,z), ghost offset 3 */
static inline int64_t orc_id(const orc *o, int x, int y, int z)
{
    int64_t id = (int64_t)(x + ORC_GHOST) + o->dy * (int64_t)(y + ORC_GHOST);
    if (o->dim == 3) id += o->dz * (int64_t)(z + ORC_GHOST);
    return id;
}

static inline int orc_wrap(int c, int L)
{
    return ((c % L) + L) % L;
}

/* flat id of the periodic image (period L = n+1 per axis) of any cell */
static inline int64_t orc_image(const orc *o, int64_t id)
{
    int64_t W = o->dy, H = o->L[1] + 2 * ORC_GHOST;
    int x = (int)(id % W) - ORC_GHOST;
    int y = (int)((id / W) % H) - ORC_GHOST;
    int z = o->dim == 3 ? (int)(id / o->dz) - ORC_GHOST : 0;
    return orc_id(o, orc_wrap(x, o->L[0]), orc_wrap(y, o->L[1]),
                  o->dim == 3 ? orc_wrap(z, o->L[2]) : 0);
}

static inline int orc_is_ghost(const orc *o, int64_t id)
{
    int64_t W = o->dy, H = o->L[1] + 2 * ORC_GHOST;
    int x = (int)(id % W) - ORC_GHOST;
    int y = (int)((id / W) % H) - ORC_GHOST;
    if (x < 0 || x >= o->L[0] || y < 0 || y >= o->L[1]) return 1;
    if (o->dim == 3) {
        int z = (int)(id / o->dz) - ORC_GHOST;
        if (z < 0 || z >= o->L[2]) return 1;
    }
    return 0;
}

/* ------------------------------------------------------------------------
 * a5  PassInt at one rank (cart_mpi.f90:281-317): every neighbour is this
 * rank, the send sub-boxes are interior and the receive sub-boxes are the
 * ghost shells (cart_mpi.f90:180-183), so every ghost cell, corners
 * included, receives its periodic image.  Order-free.
 * a10 PassF at one rank (cart_mpi.f90:240-277): the same fill for all nq
 * populations of every ghost cell.
 * Pinned by: the closed-form periodic image, checked cell by cell.
 * ---------------------------------------------------------------------- */
void orc_pass_int(orc *o, int32_t *a)
{
    for (int64_t id = 0; id < o->A; ++id)
        if (orc_is_ghost(o, id)) a[id] = a[orc_image(o, id)];
}

void orc_pass_f(orc *o, double *a)
{
    const int64_t nq = o->nq;
#pragma omp parallel for schedule(static)
    for (int64_t id = 0; id < o->A; ++id)
        if (orc_is_ghost(o, id)) {
            int64_t s = orc_image(o, id);
            for (int64_t iq = 0; iq < nq; ++iq) a[id * nq + iq] = a[s * nq + iq];
        }
}

/* ------------------------------------------------------------------------
 * a4  SetGeometry (initialization.f90:167-255; initialization_3d.f90:170-279)
 * Scan y outer / x inner (3D: z, y, x) over [-1, L+1] per axis.  Interior
 * node is fluid iff (x-pos)^2 + (y-ny/2)^2 [+ (z-nz/2)^2] > radius^2, fluid
 * nodes appended to fluid_id in scan order.  Ghost-ring cells feed the wall
 * lists: x=-1 -> bl, x=nx+1 -> br, y=-1 -> bd, y=ny+1 -> bu (3D: y -> bf/bb,
 * z -> bd/bu).  Then PassInt(geo), then the body scan that keeps, per x-row,
 * the first solid cell whose cell at offset e_x(q) + dy*(q*dim+1)
 * [+ dz*(q*dim+2)] is fluid for some q (the row is left after the first
 * hit; the offsets use q itself, initialization.f90:240, _3d:263).
 * Pinned by: fluid/solid/b_user/link counts of five inputs derived by an
 * independent analysis restatement, the circle/sphere test on every node,
 * and structural invariants of the lists (tests/test_oracle_setup.py).
 * ---------------------------------------------------------------------- */
void orc_set_geometry_shape(orc *o, double pos, double radius, double aspect, double angle_deg);
void orc_set_geometry(orc *o, double pos, double radius) { orc_set_geometry_shape(o, pos, radius, 1.0, 0.0); }

/* The same with an elliptic body (2D, no counterpart in the reference): the
 * ellipse of the circle's area with axis ratio `aspect`, axis turned by
 * angle_deg from x; a node is fluid when (u/A)^2 + (v/B)^2 > 1 in the turned
 * frame, A = radius*sqrt(aspect), B = radius/sqrt(aspect).  aspect 1,
 * angle 0 is the reference's circle test itself.  Pinned by closed forms:
 * area, symmetry, aspect 1 == circle, (aspect, 90 deg) == (1/aspect, 0). */
void orc_set_geometry_shape(orc *o, double pos, double radius, double aspect, double angle_deg)
{
    const int dim = o->dim;
    const int ellipse = dim == 2 && (aspect != 1.0 || angle_deg != 0.0);
    const double th = angle_deg * (3.14159265358979323846 / 180.0);
    const double ct = cos(th), st = sin(th);
    const double semi_a2 = radius * radius * aspect, semi_b2 = radius * radius / aspect;
    const int nx = o->n[0], ny = o->n[1], nz = o->n[2];
    o->pos = pos;
    o->radius = radius;
    int64_t idn = 0, idl = 0, idr = 0, idu = 0, idd = 0, idf = 0, idb = 0;
    int kmax = dim == 3 ? o->L[2] + 2 : 0;
    for (int k = 0; k <= kmax; ++k)
        for (int j = 0; j <= o->L[1] + 2; ++j)
            for (int i = 0; i <= o->L[0] + 2; ++i) {
                int x = i - 1, y = j - 1, z = dim == 3 ? k - 1 : 0;
                int64_t id = orc_id(o, x, y, z);
                int inside = x >= 0 && x < o->L[0] && y >= 0 && y < o->L[1] &&
                             (dim == 2 || (z >= 0 && z < o->L[2]));
                if (inside) {
                    o->geo[id] = 0;
                    double a = (double)x - pos;
                    double b = (double)y - (double)ny / 2.0;
                    double s = a * a + b * b;
                    if (dim == 3) {
                        double c = (double)z - (double)nz / 2.0;
                        s = s + c * c;
                    }
                    int fluid = s > radius * radius;
                    if (ellipse) {
                        double u = ct * a + st * b, v = ct * b - st * a;
                        fluid = u * u / semi_a2 + v * v / semi_b2 > 1.0;
                    }
                    if (fluid) {
                        o->geo[id] = 1;
                        o->fluid_id[idn++] = id;
                    }
                } else {
                    if (x == -1) o->bl[idl++] = id;
                    if (x == nx + 1) o->br[idr++] = id;
                    if (dim == 2) {
                        if (y == -1) o->bd[idd++] = id;
                        if (y == ny + 1) o->bu[idu++] = id;
                    } else {
                        if (y == -1) o->bf[idf++] = id;
                        if (y == ny + 1) o->bb[idb++] = id;
                        if (z == -1) o->bd[idd++] = id;
                        if (z == nz + 1) o->bu[idu++] = id;
                    }
                }
            }
    orc_pass_int(o, o->geo);
    o->size_fluid = idn; o->l_size = idl; o->r_size = idr;
    o->u_size = idu; o->d_size = idd; o->f_size = idf; o->b_size = idb;

    int64_t id_user = 0;
    for (int k = 0; k <= kmax; ++k)
        for (int j = 0; j <= o->L[1] + 2; ++j)
            for (int i = 0; i <= o->L[0] + 2; ++i) {
                int x = i - 1, y = j - 1, z = dim == 3 ? k - 1 : 0;
                int64_t id = orc_id(o, x, y, z);
                int hit = 0;
                for (int iq = 0; iq < o->nq; ++iq) {
                    if (o->geo[id] != 0) continue;
                    int64_t nid = id + o->e[iq * dim] + o->dy * (int64_t)(iq * dim + 1);
                    if (dim == 3) nid += o->dz * (int64_t)(iq * dim + 2);
                    if (nid < 0 || nid >= o->A) { o->ub_reads++; continue; }
                    if (o->geo[nid] == 1) {
                        o->b_user[id_user++] = id;
                        hit = 1;
                        break;
                    }
                }
                if (hit) break;
            }
    o->user_size = id_user;
}

/* ------------------------------------------------------------------------
 * a6  InitUP (initialization.f90:259-282; initialization_3d.f90:282-308)
 * r_vec(0:99) comes from the caller (the reference draws it with the
 * Fortran runtime's random_number).  Interior nodes only.
 * 2D: u_x = uu*4*y*(ny-y)/ny^2*geo*r    3D: u_x = uu*4*z*(nz-z)/nz^2*r*geo
 * Pinned by: closed-form values at chosen nodes (tests).
 * ---------------------------------------------------------------------- */
void orc_init_up(orc *o, const double *r_vec)
{
    const int dim = o->dim;
    const int ny = o->n[1], nz = o->n[2];
    int kmax = dim == 3 ? o->L[2] : 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < kmax; ++k)
        for (int j = 0; j < o->L[1]; ++j)
            for (int i = 0; i < o->L[0]; ++i) {
                int64_t id = orc_id(o, i, j, k);
                double r = 1.0 + (r_vec[id % 100] - 0.5) * 2.0 * 0.2;
                o->p[id] = 0.0;
                if (dim == 2) {
                    double y = (double)j;
                    o->u[id * 2] = o->uu * 4.0 * y * ((double)ny - y) / (double)(ny * ny) *
                                   (double)o->geo[id] * r;
                    o->u[id * 2 + 1] = 0.0;
                } else {
                    double z = (double)k;
                    o->u[id * 3] = o->uu * 4.0 * z * ((double)nz - z) / (double)(nz * nz) * r *
                                   (double)o->geo[id];
                    o->u[id * 3 + 1] = 0.0;
                    o->u[id * 3 + 2] = 0.0;
                }
            }
}

/* equilibrium of the incompressible He-Luo model, lbm.f90:70 */
static inline double orc_feq(const orc *o, int iq, double p, const double *u, double udu)
{
    double edu = 0.0;
    for (int d = 0; d < o->dim; ++d) edu = edu + (double)o->e[iq * o->dim + d] * u[d];
    return o->t[iq] * (p * 3.0 + o->rho0 * (3.0 * edu + 4.5 * (edu * edu) - 1.5 * udu));
}

/* ------------------------------------------------------------------------
 * a7  InitPDF (lbm.f90:24-48): f = feq(p,u) on fluid nodes.
 * Pinned by: sum_q feq = 3p and sum_q feq e = rho0 u (exact identities of
 * the lattice weights, checked to rounding) and closed-form values.
 * ---------------------------------------------------------------------- */
void orc_init_pdf(orc *o)
{
    const int dim = o->dim, nq = o->nq;
#pragma omp parallel for schedule(static)
    for (int64_t idn = 0; idn < o->size_fluid; ++idn) {
        int64_t id = o->fluid_id[idn];
        const double *u = &o->u[id * dim];
        double udu = 0.0;
        for (int d = 0; d < dim; ++d) udu = udu + u[d] * u[d];
        for (int iq = 0; iq < nq; ++iq) o->f[id * nq + iq] = orc_feq(o, iq, o->p[id], u, udu);
    }
}

/* ------------------------------------------------------------------------
 * a8  Collision (lbm.f90:52-76): fb = (1-ome) f + ome feq on fluid nodes.
 * Pinned by: conservation of sum f and sum f e when (p,u) are the moments
 * of f, the equilibrium fixed point, and the Taylor-Green decay rate
 * nu = (1/ome - 1/2)/3 in a periodic box (tests/test_oracle_physics.py).
 * ---------------------------------------------------------------------- */
void orc_collision(orc *o)
{
    const int dim = o->dim, nq = o->nq;
    const double ome = o->ome;
#pragma omp parallel for schedule(static)
    for (int64_t idn = 0; idn < o->size_fluid; ++idn) {
        int64_t id = o->fluid_id[idn];
        const double *u = &o->u[id * dim];
        double udu = 0.0;
        for (int d = 0; d < dim; ++d) udu = udu + u[d] * u[d];
        for (int iq = 0; iq < nq; ++iq) {
            double feq = orc_feq(o, iq, o->p[id], u, udu);
            o->fb[id * nq + iq] = (1.0 - ome) * o->f[id * nq + iq] + ome * feq;
        }
    }
}

/* fb[id*nq+iq] = fb[src*nq+qs], reading 0 (and counting) outside the array */
static inline void orc_copy(orc *o, int64_t id, int iq, int64_t src, int qs)
{
    const int64_t nq = o->nq;
    if (src < 0 || src >= o->A) {
        o->fb[id * nq + iq] = 0.0;
        o->ub_reads++;
        return;
    }
    o->fb[id * nq + iq] = o->fb[src * nq + qs];
}

static inline int64_t orc_eoff(const orc *o, int iq, int mult)
{
    int64_t off = (int64_t)mult * o->e[iq * o->dim] + (int64_t)mult * o->dy * o->e[iq * o->dim + 1];
    if (o->dim == 3) off += (int64_t)mult * o->dz * o->e[iq * o->dim + 2];
    return off;
}

/* mirror rule fb(g,q) = fb(g + 2 e_q, nq-1-q) over one index list, in order */
static void orc_mirror(orc *o, const int64_t *list, int64_t n)
{
    for (int64_t idn = 0; idn < n; ++idn) {
        int64_t id = list[idn];
        for (int iq = 0; iq < o->nq; ++iq) orc_copy(o, id, iq, id + orc_eoff(o, iq, 2), o->nq - 1 - iq);
    }
}

/* ------------------------------------------------------------------------
 * a9  BoundaryCondition (lbm.f90:96-155; lbm_3d.f90:96-157), in this order:
 *   PassF(fb); Up wall (bu) mirror; Down wall (bd) mirror; User (b_user)
 *   mirror; Inlet (bl) fb(g,q) = fb(g+1,q); Outlet (br) fb(g,q) = fb(g+e_q,q).
 * Each list is walked sequentially in list order, q inner, so a phase sees
 * the values its own earlier iterations wrote.
 * Pinned by: hand-derived known answers for single-slot probes at walls,
 * corners, inlet, outlet and the body (tests/test_oracle_bc.py).
 * ---------------------------------------------------------------------- */
void orc_boundary_condition(orc *o)
{
    orc_pass_f(o, o->fb);
    orc_mirror(o, o->bu, o->u_size);
    orc_mirror(o, o->bd, o->d_size);
    orc_mirror(o, o->b_user, o->user_size);
    for (int64_t idn = 0; idn < o->l_size; ++idn) {
        int64_t id = o->bl[idn];
        for (int iq = 0; iq < o->nq; ++iq) orc_copy(o, id, iq, id + 1, iq);
    }
    for (int64_t idn = 0; idn < o->r_size; ++idn) {
        int64_t id = o->br[idn];
        for (int iq = 0; iq < o->nq; ++iq) orc_copy(o, id, iq, id + orc_eoff(o, iq, 1), iq);
    }
}

/* ------------------------------------------------------------------------
 * a11 Propagation (lbm.f90:80-92): pull, f[id,q] = fb[id - e_q, q].
 * Pinned by: single-population delta probes landing at id + e_q.
 * ---------------------------------------------------------------------- */
void orc_propagation(orc *o)
{
    const int64_t nq = o->nq;
#pragma omp parallel for schedule(static)
    for (int64_t idn = 0; idn < o->size_fluid; ++idn) {
        int64_t id = o->fluid_id[idn];
        for (int iq = 0; iq < nq; ++iq) o->f[id * nq + iq] = o->fb[(id - orc_eoff(o, iq, 1)) * nq + iq];
    }
}

/* ------------------------------------------------------------------------
 * a12 PostProcessing (lbm.f90:159-201; lbm_3d.f90:161-203)
 * p = (sum_q f)/3*geo, u_d = (sum_q f e_d)/rho0*geo, sums in q order; then
 * inlet u at x=0 (all rows of bl, 2D profile in y, 3D profile in z with u_z
 * left as computed), then outlet p(nx,.) = 0.
 * Pinned by: closed-form moments of hand-built populations.
 * ---------------------------------------------------------------------- */
void orc_moments(orc *o)
{
    const int dim = o->dim, nq = o->nq;
#pragma omp parallel for schedule(static)
    for (int64_t idn = 0; idn < o->size_fluid; ++idn) {
        int64_t id = o->fluid_id[idn];
        double p = 0.0, u[3] = {0.0, 0.0, 0.0};
        for (int iq = 0; iq < nq; ++iq) {
            double fq = o->f[id * nq + iq];
            p = p + fq;
            for (int d = 0; d < dim; ++d) u[d] = u[d] + fq * (double)o->e[iq * dim + d];
        }
        for (int d = 0; d < dim; ++d) o->u[id * dim + d] = u[d] / o->rho0 * (double)o->geo[id];
        o->p[id] = p / 3.0 * (double)o->geo[id];
    }
}

/* the inlet and outlet conditions PostProcessing applies after the moments */
void orc_post_bc(orc *o)
{
    const int dim = o->dim;
    const int ny = o->n[1], nz = o->n[2];
    for (int64_t idn = 0; idn < o->l_size; ++idn) {
        int64_t id = o->bl[idn] + 1;
        if (dim == 2) {
            int y = (int)(id / o->dy) - ORC_GHOST;
            o->u[id * 2] = o->uu * o->rho0 * 4.0 * (double)y * (double)(ny - y) / (double)(ny * ny);
            o->u[id * 2 + 1] = 0.0;
        } else {
            int z = (int)(id / o->dz) - ORC_GHOST;
            o->u[id * 3] = o->uu * o->rho0 * 4.0 * (double)z * (double)(nz - z) / (double)(nz * nz);
            o->u[id * 3 + 1] = 0.0;
        }
    }
    for (int64_t idn = 0; idn < o->r_size; ++idn) o->p[o->br[idn] - 1] = 0.0;
}

void orc_post_processing(orc *o)
{
    orc_moments(o);
    orc_post_bc(o);
}

/* ------------------------------------------------------------------------
 * a13 Monitor local maximum (lbm.f90:205-217): max over fluid entries
 * idn = 1..size_fluid-1 of sqrt(u_x^2 + u_y^2) (u_z ignored in 3D).
 * Pinned by: brute-force maximum over the velocity field.
 * ---------------------------------------------------------------------- */
double orc_monitor(orc *o)
{
    double um = 0.0;
    const int dim = o->dim;
    for (int64_t idn = 1; idn < o->size_fluid; ++idn) {
        int64_t id = o->fluid_id[idn];
        double v = sqrt(o->u[id * dim] * o->u[id * dim] + o->u[id * dim + 1] * o->u[id * dim + 1]);
        if (v > um) um = v;
    }
    return um;
}

/* ------------------------------------------------------------------------
 * a14 Force on the body by momentum exchange (no counterpart in the
 * reference, which logs no force).  Call after BoundaryCondition: for every
 * fluid node x and direction q whose neighbour x+e_q is an interior solid
 * cell, the population fb(x,q) leaves toward the body and fb(x+e_q, nq-1-q)
 * -- the mirror value on a b_user cell, 0 on any other solid cell -- comes
 * back:  F += e_q (fb(x,q) + fb(x+e_q, nq-1-q)).  Summed in fluid_id order,
 * q inner, in FP64.
 * Pinned by: F = 0 at rest, F_y = 0 for a flow symmetric about the body's
 * centre line, and the sign of F_x for a uniform stream (tests).
 * ---------------------------------------------------------------------- */
void orc_forces(orc *o, double F[3])
{
    const int dim = o->dim, nq = o->nq;
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t idn = 0; idn < o->size_fluid; ++idn) {
        int64_t id = o->fluid_id[idn];
        for (int iq = 0; iq < nq; ++iq) {
            int64_t nb = id + orc_eoff(o, iq, 1);
            if (o->geo[nb] != 0 || orc_is_ghost(o, nb)) continue;
            double m = o->fb[id * nq + iq] + o->fb[nb * nq + (nq - 1 - iq)];
            for (int d = 0; d < dim; ++d) s[d] = s[d] + (double)o->e[iq * dim + d] * m;
        }
    }
    F[0] = s[0]; F[1] = s[1]; F[2] = dim == 3 ? s[2] : 0.0;
}

/* ------------------------------------------------------------------------
 * a15 Loop orchestration (main.f90:48-63) for iterations iter0..iter1
 * inclusive: Monitor when mod(iter,interv)==0, then Collision,
 * BoundaryCondition, Propagation, PostProcessing.  Monitor values are
 * stored in um_out (capacity cap); returns the number stored.
 * ---------------------------------------------------------------------- */
int64_t orc_run(orc *o, int64_t iter0, int64_t iter1, int64_t interv, double *um_out, int64_t cap)
{
    int64_t k = 0;
    for (int64_t iter = iter0; iter <= iter1; ++iter) {
        if (interv > 0 && iter % interv == 0) {
            double um = orc_monitor(o);
            if (um_out && k < cap) um_out[k] = um;
            ++k;
        }
        orc_collision(o);
        orc_boundary_condition(o);
        orc_propagation(o);
        orc_post_processing(o);
    }
    return k;
}

/* ---------------- accessors for the ctypes binding ---------------- */
double *orc_ptr_f(orc *o) { return o->f; }
double *orc_ptr_fb(orc *o) { return o->fb; }
double *orc_ptr_p(orc *o) { return o->p; }
double *orc_ptr_u(orc *o) { return o->u; }
int32_t *orc_ptr_geo(orc *o) { return o->geo; }
int64_t *orc_ptr_list(orc *o, int which)
{
    switch (which) {
    case 0: return o->fluid_id;
    case 1: return o->bl;
    case 2: return o->br;
    case 3: return o->bu;
    case 4: return o->bd;
    case 5: return o->bf;
    case 6: return o->bb;
    case 7: return o->b_user;
    default: return NULL;
    }
}
int64_t orc_list_size(orc *o, int which)
{
    switch (which) {
    case 0: return o->size_fluid;
    case 1: return o->l_size;
    case 2: return o->r_size;
    case 3: return o->u_size;
    case 4: return o->d_size;
    case 5: return o->f_size;
    case 6: return o->b_size;
    case 7: return o->user_size;
    default: return -1;
    }
}
int64_t orc_array_size(orc *o) { return o->A; }
int64_t orc_ub_reads(orc *o) { return o->ub_reads; }
void orc_reset_ub_reads(orc *o) { o->ub_reads = 0; }
const int *orc_ptr_e(orc *o) { return o->e; }
const double *orc_ptr_t(orc *o) { return o->t; }
