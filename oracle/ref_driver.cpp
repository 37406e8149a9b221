// ref_driver.cpp -- C entry points into the compiled reference sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile (target `ref`) against
// the reference sources where they lie (NgFEP/NexusMD, compiled in place,
// never copied) into oracle/_ref/libnxref.so.  It is used to pin the C oracle
// (nxo.c) and to write the golden fixtures under tests/golden/.
//
// The reference's NonbondedForce::perform3DFFT calls FFTW 3.3.5, which exists
// in the reference only as a Win64 DLL.  This driver therefore never reaches
// perform3DFFT, ReciprocalPMEcalculateForcesAndEnergy, calculateForces or
// Forces::AddNonBondElectroPME; the link drops them (-ffunction-sections,
// --gc-sections) and -z defs proves no FFTW symbol is left unresolved.  The
// PME pieces either side of the FFT (spread, convolution, gather, B-spline
// setup) are driven individually, with the private members reached through
// `#define private public` at include time.
#include <complex>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "Coords3D.h"
#include "NeighborList.h"
#include "Types.h"
#include "PeriodicBoundaryCondition.h"
#include "SystemXMLParser.h"
#include "StateXMLParser.h"
#define private public
#include "NonbondedForce.h"
#include "Reporter.h"
#undef private
#include "Exclusions.h"
#include "Forces.h"
#include "KineticEnergy.h"
#include "NeighborList.h"
#include "VerletIntegration.h"

using namespace BaseLine;

namespace {
std::vector<Coords3D> to_coords(int n, const double* p) {
    std::vector<Coords3D> v(n);
    for (int i = 0; i < n; ++i) v[i] = Coords3D(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    return v;
}
void from_coords(const std::vector<Coords3D>& v, double* p) {
    for (size_t i = 0; i < v.size(); ++i)
        for (int k = 0; k < 3; ++k) p[3 * i + k] = v[i][k];
}
PeriodicBoundaryCondition::BoxInfo make_box(const double* lb, const double* ub) {
    return PeriodicBoundaryCondition::BoxInfo(Coords3D(lb[0], lb[1], lb[2]), Coords3D(ub[0], ub[1], ub[2]));
}
std::vector<std::set<int>> excl_sets(int n, const int* off, const int* idx) {
    std::vector<std::set<int>> ex(n);
    if (off)
        for (int i = 0; i < n; ++i)
            for (int a = off[i]; a < off[i + 1]; ++a) ex[i].insert(idx[a]);
    return ex;
}
NonbondedParams nb_params(int n, const double* q, const double* sig, const double* eps, double cutoff,
                          double tol, int use_switch, double rsw) {
    NonbondedParams p{};
    p.cutoff = cutoff;
    p.ewaldTolerance = tol;
    p.useSwitchingFunction = use_switch;
    p.switchingDistance = rsw;
    p.particles.resize(n);
    for (int i = 0; i < n; ++i) p.particles[i] = NonbondedParticle{q[i], sig[i], eps[i]};
    return p;
}
}  // namespace

#pragma GCC visibility push(default)
extern "C" {

// PeriodicBoundaryCondition::extractBoxBoundaries
void nxr_box_from_positions(int n, const double* pos, double* lb, double* ub, double* L) {
    PeriodicBoundaryCondition::BoxInfo b;
    auto v = to_coords(n, pos);
    PeriodicBoundaryCondition::extractBoxBoundaries(v, b, "");
    for (int k = 0; k < 3; ++k) { lb[k] = b.lb[k]; ub[k] = b.ub[k]; L[k] = b.boxSize[k]; }
}

// Exclusions::createExclusions; returns the total entry count, fills off/idx if idx_cap suffices.
long nxr_exclusions(int n, int nb, const int* p1, const int* p2, int* off, int* idx, long idx_cap) {
    std::vector<BondParams> bonds(nb);
    for (int b = 0; b < nb; ++b) bonds[b] = BondParams{p1[b], p2[b], 0.0, 0.0};
    std::vector<std::set<int>> ex(n);
    int nn = n, cut = 3;
    Exclusions::createExclusions(nn, bonds, ex, cut);
    long tot = 0;
    off[0] = 0;
    for (int i = 0; i < n; ++i) {
        for (int j : ex[i]) { if (tot < idx_cap) idx[tot] = j; ++tot; }
        off[i + 1] = (int)tot;
    }
    return tot;
}

// obtainNeighborList (pairs as emitted: first > second)
long nxr_pairs(int n, const double* pos, const double* lb, const double* ub, const int* off, const int* idx,
               double cutoff, int* pairs, long cap) {
    auto v = to_coords(n, pos);
    auto box = make_box(lb, ub);
    auto ex = excl_sets(n, off, idx);
    NeighborList nl;
    obtainNeighborList(nl, v, box, ex, cutoff, 0.0, true);
    long np = (long)nl.pairs.size();
    for (long p = 0; p < np && p < cap; ++p) { pairs[2 * p] = nl.pairs[p].first; pairs[2 * p + 1] = nl.pairs[p].second; }
    return np;
}

// Forces::AddHBond / AddHAngle / AddPTorsion
double nxr_bonds(int n, const double* pos, const double* lb, const double* ub, int nb, const int* p1,
                 const int* p2, const double* d, const double* k, double* f) {
    std::vector<BondParams> b(nb);
    for (int t = 0; t < nb; ++t) b[t] = BondParams{p1[t], p2[t], d[t], k[t]};
    auto v = to_coords(n, pos);
    std::vector<Coords3D> F(n, Coords3D(0, 0, 0));
    double e = 0.0;
    Forces::AddHBond(F, v, b, e, make_box(lb, ub));
    from_coords(F, f);
    return e;
}

double nxr_angles(int n, const double* pos, const double* lb, const double* ub, int na, const int* p1,
                  const int* p2, const int* p3, const double* a, const double* k, double* f) {
    std::vector<AngleParams> A(na);
    for (int t = 0; t < na; ++t) A[t] = AngleParams{p1[t], p2[t], p3[t], a[t], k[t]};
    auto v = to_coords(n, pos);
    std::vector<Coords3D> F(n, Coords3D(0, 0, 0));
    double e = 0.0;
    Forces::AddHAngle(F, v, A, e, make_box(lb, ub));
    from_coords(F, f);
    return e;
}

double nxr_torsions(int n, const double* pos, const double* lb, const double* ub, int nt, const int* p1,
                    const int* p2, const int* p3, const int* p4, const double* k, const double* phase,
                    const int* per, double* f) {
    std::vector<PTorsionParams> T(nt);
    for (int t = 0; t < nt; ++t) T[t] = PTorsionParams{p1[t], p2[t], p3[t], p4[t], k[t], phase[t], per[t]};
    auto v = to_coords(n, pos);
    std::vector<Coords3D> F(n, Coords3D(0, 0, 0));
    double e = 0.0;
    Forces::AddPTorsion(F, v, T, e, make_box(lb, ub));
    from_coords(F, f);
    return e;
}

// NonbondedForce::InitializeAlphaEwald and initializeGridSize
double nxr_alpha(double cutoff, double tol) {
    NonbondedForce nb;
    NonbondedParams p{};
    p.cutoff = cutoff;
    p.ewaldTolerance = tol;
    nb.InitializeAlphaEwald(p);
    return nb.alphaEwald;
}

void nxr_grid_rule(const double* lb, const double* ub, int* grid) {
    NonbondedForce nb;
    nb.initializeGridSize(make_box(lb, ub));
    for (int d = 0; d < 3; ++d) grid[d] = nb.gridSize[d];
}

// computeBSplineParameters for a given grid: moduli[3][K], theta/dtheta[3][n][5], index[n][3], frac[n][3]
void nxr_bspline(int n, const double* pos, const double* lb, const double* ub, const int* grid, double* mod0,
                 double* mod1, double* mod2, double* theta, double* dtheta, int* index, double* frac) {
    NonbondedForce nb;
    auto box = make_box(lb, ub);
    auto v = to_coords(n, pos);
    nb.initializeReciprocalVectors(box);
    nb.gridSize = {grid[0], grid[1], grid[2]};
    nb.InitparticleIndexAndFraction(v);
    nb.computeBSplineParameters(5, v);
    double* mods[3] = {mod0, mod1, mod2};
    for (int d = 0; d < 3; ++d)
        for (int k = 0; k < grid[d]; ++k) mods[d][k] = nb.bsplines_moduli[d][k];
    for (int d = 0; d < 3; ++d)
        for (int i = 0; i < n; ++i)
            for (int o = 0; o < 5; ++o) {
                theta[(d * n + i) * 5 + o] = nb.bsplines_theta[d][i][o];
                dtheta[(d * n + i) * 5 + o] = nb.bsplines_dtheta[d][i][o];
            }
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) { index[3 * i + d] = nb.particleIndex[i][d]; frac[3 * i + d] = nb.particleFraction[i][d]; }
}

// gridSpreadCharge_Reference into an interleaved complex grid
void nxr_pme_spread(int n, const double* pos, const double* q, const double* lb, const double* ub, const int* grid,
                    double* cgrid) {
    NonbondedForce nb;
    auto box = make_box(lb, ub);
    auto v = to_coords(n, pos);
    nb.initializeReciprocalVectors(box);
    nb.gridSize = {grid[0], grid[1], grid[2]};
    nb.InitparticleIndexAndFraction(v);
    nb.computeBSplineParameters(5, v);
    NonbondedParams p = nb_params(n, q, q, q, 1.0, 5e-4, 0, 0.0);
    ComplexGrid g((size_t)grid[0] * grid[1] * grid[2]);
    nb.gridSpreadCharge_Reference(v, g, p, box, 5);
    std::memcpy(cgrid, g.data(), sizeof(double) * 2 * g.size());
}

// reciprocalConvolutionAndEnergy on a caller grid (in place); returns the energy it assigns
double nxr_pme_convolve(const double* lb, const double* ub, const int* grid, double cutoff, double tol,
                        double* cgrid) {
    NonbondedForce nb;
    auto box = make_box(lb, ub);
    std::vector<Coords3D> none;
    nb.initializeReciprocalVectors(box);
    nb.gridSize = {grid[0], grid[1], grid[2]};
    nb.InitparticleIndexAndFraction(none);
    nb.computeBSplineParameters(5, none);
    NonbondedParams p{};
    p.cutoff = cutoff;
    p.ewaldTolerance = tol;
    nb.InitializeAlphaEwald(p);
    size_t G = (size_t)grid[0] * grid[1] * grid[2];
    ComplexGrid g(G);
    std::memcpy(g.data(), cgrid, sizeof(double) * 2 * G);
    double e = 0.0;
    nb.reciprocalConvolutionAndEnergy(g, p, box, e);
    std::memcpy(cgrid, g.data(), sizeof(double) * 2 * G);
    return e;
}

// PMEGridForces from a caller grid (forces accumulated into f with -=)
void nxr_pme_gather(int n, const double* pos, const double* q, const double* lb, const double* ub, const int* grid,
                    const double* cgrid, double* f) {
    NonbondedForce nb;
    auto box = make_box(lb, ub);
    auto v = to_coords(n, pos);
    nb.initializeReciprocalVectors(box);
    nb.gridSize = {grid[0], grid[1], grid[2]};
    nb.InitparticleIndexAndFraction(v);
    nb.computeBSplineParameters(5, v);
    NonbondedParams p = nb_params(n, q, q, q, 1.0, 5e-4, 0, 0.0);
    size_t G = (size_t)grid[0] * grid[1] * grid[2];
    ComplexGrid g(G);
    std::memcpy(g.data(), cgrid, sizeof(double) * 2 * G);
    std::vector<Coords3D> F = to_coords(n, f);
    double pe = 0.0;
    nb.PMEGridForces(F, v, g, p, box, pe, 5);
    from_coords(F, f);
}

// NonbondedParamsModifier + CalculateSelfEnergy + DirectForcesAndEnergy + CalculateExclusionEnergyAndForces
// e_out[0] = self, e_out[1] = direct (coul + LJ), e_out[2] = -exclusion (as added to the PE)
void nxr_nonbonded_real(int n, const double* pos, const double* q, const double* sig, const double* eps,
                        const double* lb, const double* ub, const int* off, const int* idx, double cutoff,
                        double tol, int use_switch, double rsw, double* f, double* e_out) {
    NonbondedForce nb;
    auto box = make_box(lb, ub);
    auto v = to_coords(n, pos);
    auto ex = excl_sets(n, off, idx);
    NonbondedParams p = nb_params(n, q, sig, eps, cutoff, tol, use_switch, rsw);
    nb.InitializeAlphaEwald(p);
    nb.NonbondedParamsModifier(v, p);
    double e = 0.0;
    nb.CalculateSelfEnergy(p, e);
    e_out[0] = e;
    std::vector<Coords3D> F(n, Coords3D(0, 0, 0));
    e = 0.0;
    nb.DirectForcesAndEnergy(F, v, p, box, ex, e);
    e_out[1] = e;
    e = 0.0;
    nb.CalculateExclusionEnergyAndForces(F, v, box, ex, e);
    e_out[2] = e;
    from_coords(F, f);
}

// VerletIntegration::InverseMasses + Advance (dt0 = dt1 = dt)
void nxr_leapfrog(int n, double* pos, double* vel, const double* f, const double* mass, double dt, const double* lb,
                  const double* ub) {
    auto P = to_coords(n, pos), V = to_coords(n, vel), F = to_coords(n, f);
    std::vector<double> m(mass, mass + n), im;
    VerletIntegration::InverseMasses(m, im);
    int step = 0;
    std::vector<double> dts = {dt, dt};
    VerletIntegration::Advance(P, V, F, im, step, dts, make_box(lb, ub));
    from_coords(P, pos);
    from_coords(V, vel);
}

// KineticEnergy::calculateKineticEnergy
double nxr_kinetic(int n, const double* vel, const double* mass, const double* f, double dt) {
    auto V = to_coords(n, vel), F = to_coords(n, f);
    std::vector<double> m(mass, mass + n), ke(n, 0.0);
    double tot = 0.0;
    KineticEnergy::calculateKineticEnergy(V, m, F, dt, n, ke, tot);
    return tot;
}

// SystemXMLParser / StateXMLParser: counts[0..4] = masses, bonds, angles, torsions, particles
void nxr_parse_counts(const char* system_xml, const char* state_xml, int* counts) {
    counts[0] = (int)SystemXMLParser::MassesParser(system_xml).size();
    counts[1] = (int)SystemXMLParser::HBondParser(system_xml).size();
    counts[2] = (int)SystemXMLParser::HAngleParser(system_xml).size();
    counts[3] = (int)SystemXMLParser::PTorsionParser(system_xml).size();
    counts[4] = (int)SystemXMLParser::NonBondedParser(system_xml).particles.size();
    counts[5] = state_xml ? (int)StateXMLParser::parseStateXML(state_xml).size() : 0;
}

// Reporter byte formats (writes files relative to the CWD, like the reference)
void nxr_report(const char* out_name, int n, const double* pos, const double* vel, const double* f, double ke,
                double pe, int step) {
    Reporter r;
    auto P = to_coords(n, pos), V = to_coords(n, vel), F = to_coords(n, f);
    double te = pe + ke;
    r.TotalEnergyReport(out_name, ke, pe, te, step);
    r.TestPVFReport(out_name, P, V, F, step, {}, {}, {});
}

void nxr_pdb_report(const char* in_pdb, const char* out_pdb, int n, const double* pos, int model, int first) {
    Reporter r;
    std::vector<PDBAtom> tmpl;
    if (first) r.pdbOutputGeneratorPart1(in_pdb, out_pdb, tmpl);
    else r.outputTemplateExtractor(in_pdb, tmpl);
    r.pdbOutputGeneratorPart2(out_pdb, tmpl, to_coords(n, pos), model);
}

}  // extern "C"
#pragma GCC visibility pop
