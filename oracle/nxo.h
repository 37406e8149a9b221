/*
 * nxo.h -- CPU FP64 oracle for the NexusMD force-and-integrate path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithms (NgFEP/NexusMD, NexusMD/core/...) used as the checker for the
 * HIP engine and as the timed CPU baseline.  Only tests/, the smoke check and
 * the bench's cpu_baseline leg may load it; the product library never links,
 * loads or calls it.
 *
 * Every function cites the reference file:line it follows.  Paths are
 * relative to the reference repository root; "core/" = "NexusMD/core/".
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   nxo_box_from_positions ..... compiled reference (oracle/_ref) + closed form
 *   nxo_exclusions ............. compiled reference + BFS graph distance
 *   nxo_pairs_voxel ............ compiled reference (bit-exact pair set)
 *   nxo_pairs_brute ............ brute force O(N^2) (mathematical set)
 *   nxo_bond/angle/torsion ..... compiled reference + finite differences
 *   nxo_direct/self/exclusion .. compiled reference + closed form + FD
 *   nxo_bspline_* / moduli ..... compiled reference + closed-form M5 values
 *   nxo_fft3d .................. numpy.fft (pocketfft) + brute DFT
 *   nxo_pme_recip .............. reference pieces chained around nxo_fft3d +
 *                                converged direct Ewald sum + Madelung NaCl
 *   nxo_leapfrog / nxo_kinetic . compiled reference + closed form
 */
#ifndef NXO_H
#define NXO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Physical constants, core/baseline/headers/Forces/NonbondedForce.h:13-28 */
#define NXO_PI 3.14159265358979323846
#define NXO_SQRT_PI 1.77245385090551602729
#define NXO_TWO_OVER_SQRT_PI 1.1283791670955126
#define NXO_ONE_4PI_EPS0 138.93545764446428693393

typedef struct { double lb[3], ub[3], L[3]; } nxo_box;

typedef struct { int p1, p2; double d, k; } nxo_bond;                       /* SystemXMLParser.h:16-21 */
typedef struct { int p1, p2, p3; double a, k; } nxo_angle;                  /* SystemXMLParser.h:23-27 */
typedef struct { int p1, p2, p3, p4; double k, phase; int periodicity; } nxo_torsion; /* :10-14 */

typedef struct {
    int n;
    const double* mass;
    int n_bonds;    const nxo_bond* bonds;
    int n_angles;   const nxo_angle* angles;
    int n_torsions; const nxo_torsion* torsions;
    const double* q; const double* sig; const double* eps;
    const int* excl_off; const int* excl_idx;     /* symmetric CSR */
    double cutoff, ewald_tol;
    int grid[3];                 /* PME grid; 0 = reference rule (next pow2 >= 10*ceil(L)) */
    int use_switch; double switch_distance;
    int switch_quirk;            /* 1 = reproduce the reference's switch carry-over between pairs */
    nxo_box box;
} nxo_system;

/* Energy terms, kJ/mol */
typedef struct {
    double bond, angle, torsion, coulomb, lj, recip, self, exclusion;
    double potential;      /* physical sum of all enabled terms */
    double reference_pe;   /* what Engine reports (recip energy assigned over the bonded sum) */
} nxo_energies;

enum { NXO_BOND = 1, NXO_ANGLE = 2, NXO_TORSION = 4, NXO_NONBONDED = 8 };

/* --- box and minimum image --- */
void nxo_box_from_positions(int n, const double* pos, nxo_box* box);
void nxo_box_from_bounds(const double lb[3], const double ub[3], nxo_box* box);
void nxo_min_image(const double* a, const double* b, const nxo_box* box, double* d);

/* --- exclusions (CSR; caller frees *off and *idx with nxo_free) --- */
int nxo_exclusions(int n, int n_bonds, const nxo_bond* bonds, int bond_cutoff,
                   int** off, int** idx);
void nxo_free(void* p);

/* --- neighbour list: pairs (i, j) with i > j, r^2 < rc^2, j not excluded --- */
int64_t nxo_pairs_voxel(int n, const double* pos, const nxo_box* box,
                        const int* excl_off, const int* excl_idx, double cutoff,
                        int* pairs, int64_t cap);
int64_t nxo_pairs_brute(int n, const double* pos, const nxo_box* box,
                        const int* excl_off, const int* excl_idx, double cutoff,
                        int* pairs, int64_t cap);

/* --- bonded terms: forces are accumulated (+=), energy returned --- */
double nxo_bonds(int n_bonds, const nxo_bond* b, const double* pos, const nxo_box* box, double* f);
double nxo_angles(int n_angles, const nxo_angle* a, const double* pos, const nxo_box* box, double* f);
double nxo_torsions(int n_torsions, const nxo_torsion* t, const double* pos, const nxo_box* box, double* f);

/* --- nonbonded pieces --- */
double nxo_alpha(double cutoff, double ewald_tol);
void   nxo_grid_rule(const nxo_box* box, int grid[3]);
void   nxo_bspline_moduli(int K, int order, double* moduli);
void   nxo_bspline_theta(double frac, int order, double* theta, double* dtheta);
void   nxo_particle_index(const double* x, const nxo_box* box, const int grid[3], int idx[3], double frac[3]);
void   nxo_fft3d(int nx, int ny, int nz, double* data /* interleaved complex */, int sign);
void   nxo_fft1d(int n, double* data, int stride_complex, int sign);
void   nxo_pme_spread(int n, const double* pos, const double* q, const nxo_box* box,
                      const int grid[3], double* cgrid /* complex, zeroed here */);
double nxo_pme_convolve(double* cgrid, const nxo_box* box, const int grid[3], double alpha);
void   nxo_pme_gather(int n, const double* pos, const double* q, const nxo_box* box,
                      const int grid[3], const double* cgrid, double* f);
double nxo_pme_recip(int n, const double* pos, const double* q, const nxo_box* box,
                     const int grid[3], double alpha, double* f);
double nxo_self_energy(int n, const double* q, double alpha);
double nxo_direct(int n, const double* pos, const double* q, const double* sig, const double* eps,
                  const nxo_box* box, const int* pairs, int64_t n_pairs, double alpha,
                  double cutoff, int use_switch, double switch_distance, int switch_quirk,
                  double* f, double* e_coul, double* e_lj);
double nxo_exclusion_correction(int n, const double* pos, const double* q, const nxo_box* box,
                                const int* excl_off, const int* excl_idx, double alpha, double* f);

/* --- whole force evaluation (Engine::CalculateForces) --- */
int nxo_compute_forces(const nxo_system* sys, const double* pos, int mask,
                       double* f, nxo_energies* e);

/* --- integration and kinetic energy --- */
void   nxo_leapfrog(int n, double* pos, double* vel, const double* f, const double* mass,
                    double dt, const nxo_box* box);
double nxo_kinetic(int n, const double* vel, const double* mass, const double* f, double dt);

/* --- Engine::RunSimulation without file output: per report step writes
 *     (KE, reference PE, total) into energies_out[3*k] --- */
int nxo_simulate(const nxo_system* sys, double* pos, double* vel, int mask,
                 int nsteps, double dt, int interval, double* energies_out, double* f_last);

#ifdef __cplusplus
}
#endif
#endif
