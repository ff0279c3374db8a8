/* oracle_check.c -- runs the oracle (oracle/imexlb_oracle.c) end to end for
 * one lattice, for builds with AddressSanitizer / UndefinedBehaviorSanitizer:
 * create, SetGeometry, InitUP, InitPDF, steps with Monitor, forces.  Prints
 * "monitor <last Umax> forces <Fx> <Fy> <Fz> ub <reads past the array>"
 * with %.17g, so a caller can compare with the uninstrumented build.
 *
 *   oracle_check dim nx ny nz pos radius ome rho0 uu aspect angle steps
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

typedef struct orc orc;
orc *orc_create(int dim, int nx, int ny, int nz);
void orc_destroy(orc *o);
void orc_set_params(orc *o, double ome, double rho0, double uu);
void orc_set_geometry_shape(orc *o, double pos, double radius, double aspect, double angle_deg);
void orc_init_up(orc *o, const double *r_vec);
void orc_init_pdf(orc *o);
double orc_monitor(orc *o);
void orc_forces(orc *o, double F[3]);
int64_t orc_run(orc *o, int64_t iter0, int64_t iter1, int64_t interv, double *um_out, int64_t cap);
int64_t orc_ub_reads(orc *o);

int main(int argc, char **argv)
{
    if (argc != 13) {
        fprintf(stderr, "usage: %s dim nx ny nz pos radius ome rho0 uu aspect angle steps\n", argv[0]);
        return 2;
    }
    const int dim = atoi(argv[1]), nx = atoi(argv[2]), ny = atoi(argv[3]), nz = atoi(argv[4]);
    const double pos = atof(argv[5]), radius = atof(argv[6]), ome = atof(argv[7]), rho0 = atof(argv[8]);
    const double uu = atof(argv[9]), aspect = atof(argv[10]), angle = atof(argv[11]);
    const int64_t steps = atoll(argv[12]);
    orc *o = orc_create(dim, nx, ny, nz);
    if (!o) return 3;
    orc_set_params(o, ome, rho0, uu);
    orc_set_geometry_shape(o, pos, radius, aspect, angle);
    double r_vec[100];
    for (int k = 0; k < 100; ++k) r_vec[k] = 0.5 + 0.45 * sin(1.7 * k + 0.3);
    orc_init_up(o, r_vec);
    orc_init_pdf(o);
    double um[64];
    const int64_t interv = steps / 4 > 0 ? steps / 4 : 1;
    const int64_t n = orc_run(o, 0, steps - 1, interv, um, 64);
    double F[3] = {0, 0, 0};
    orc_forces(o, F);
    printf("monitor %.17g forces %.17g %.17g %.17g ub %lld samples %lld\n", orc_monitor(o), F[0], F[1], F[2],
           (long long)orc_ub_reads(o), (long long)n);
    orc_destroy(o);
    return 0;
}
