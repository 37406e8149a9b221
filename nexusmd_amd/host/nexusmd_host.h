// nexusmd_host.h -- C++ host side of the MI355X engine, mirroring NexusMD's
// dispatcher interface so that a caller of the reference finds the same
// classes, names, argument meanings and file formats:
//
//   TaskDispatcher      core/dispatcher/headers/TaskDispatcher.h:9-43
//                       (CreateDispatcher, Enable*Force, Simulate)
//   SystemXMLParser     core/common/headers/SystemXMLParser.h:43-60
//   StateXMLParser      core/common/headers/StateXMLParser.h
//   Reporter            core/baseline/headers/Reporter.h (PDB, PVF and
//                       TotalEnergy byte formats, Reporter.cpp:26-230)
//
// Simulate runs on the GPU through the C-ABI of include/nexusmd.h; nothing
// here computes forces.  Types carry the reference's field names and units
// (nm, ps, amu, kJ/mol, e).
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "nexusmd.h"

namespace nexusmd {

// Three-component vector with the interface of the reference's Coords3D
// (core/common/headers/Coords3D.h:15-151): components by index, component-
// wise + and -, scaling by a double, dot and cross products, length and a
// normalised copy (the zero vector stays zero).  dot adds x, y, z left to
// right, as the reference does.
class Coords3D {
   public:
    Coords3D() = default;
    Coords3D(double x, double y, double z) : data{x, y, z} {}
    double operator[](int i) const { return data[i]; }
    double& operator[](int i) { return data[i]; }
    bool operator==(const Coords3D& o) const { return data[0] == o[0] && data[1] == o[1] && data[2] == o[2]; }
    bool operator!=(const Coords3D& o) const { return !(*this == o); }
    Coords3D operator+() const { return *this; }
    Coords3D operator-() const { return Coords3D(-data[0], -data[1], -data[2]); }
    Coords3D operator+(const Coords3D& o) const { return Coords3D(data[0] + o[0], data[1] + o[1], data[2] + o[2]); }
    Coords3D operator-(const Coords3D& o) const { return Coords3D(data[0] - o[0], data[1] - o[1], data[2] - o[2]); }
    Coords3D& operator+=(const Coords3D& o) { return *this = *this + o; }
    Coords3D& operator-=(const Coords3D& o) { return *this = *this - o; }
    Coords3D operator*(double s) const { return Coords3D(data[0] * s, data[1] * s, data[2] * s); }
    Coords3D& operator*=(double s) { return *this = *this * s; }
    // division multiplies by the reciprocal, as the reference does
    Coords3D operator/(double s) const { return *this * (1.0 / s); }
    Coords3D& operator/=(double s) { return *this = *this / s; }
    double dot(const Coords3D& o) const { return data[0] * o[0] + data[1] * o[1] + data[2] * o[2]; }
    Coords3D cross(const Coords3D& o) const {
        return Coords3D(data[1] * o[2] - data[2] * o[1], data[2] * o[0] - data[0] * o[2], data[0] * o[1] - data[1] * o[0]);
    }
    double length() const { return std::sqrt(dot(*this)); }
    // each component divided by the length
    Coords3D normalize() const {
        const double l = length();
        return l > 0.0 ? Coords3D(data[0] / l, data[1] / l, data[2] / l) : Coords3D();
    }
    double data[3] = {0.0, 0.0, 0.0};
};

using BondParams = nx_bond;          // p1, p2, d, k          (SystemXMLParser.h:16-21)
using AngleParams = nx_angle;        // p1, p2, p3, a, k      (SystemXMLParser.h:23-27)
using PTorsionParams = nx_torsion;   // p1..p4, k, phase, periodicity (SystemXMLParser.h:10-14)

struct NonbondedParticle { double q, sig, eps; };
struct NonbondedException { int p1, p2; double q, sig, eps; };

// Attributes of the first <Force type="NonbondedForce"> (SystemXMLParser.h:34-39)
struct NonbondedParams {
    double alpha = 0, cutoff = 0, ewaldTolerance = 0, rfDielectric = 0, switchingDistance = 0;
    int dispersionCorrection = 0, exceptionsUsePeriodic = 0, forceGroup = 0, includeDirectSpace = 0, ljAlpha = 0,
        ljnx = 0, ljny = 0, ljnz = 0, method = 0, nx = 0, ny = 0, nz = 0, recipForceGroup = 0,
        useSwitchingFunction = 0, version = 0;
    bool present = false;
    std::vector<NonbondedParticle> particles;
    std::vector<NonbondedException> exceptions;
};

// Everything Engine::InitializeSimulationParameters reads from the system XML
// (Engine.cpp:55-64), gathered in one pass over the file.
struct SystemDescription {
    std::vector<double> masses;
    std::vector<BondParams> bonds;
    std::vector<AngleParams> angles;
    std::vector<PTorsionParams> torsions;
    NonbondedParams nonbonded;
};

class SystemXMLParser {
   public:
    // One pass over the file; throws std::runtime_error if it cannot be read or parsed.
    static SystemDescription Parse(const std::string& filename);
    // The reference's per-section entry points (each reports on std::cerr and
    // returns an empty result on a file error, as the reference does).
    static std::vector<double> MassesParser(const std::string& filename);
    static std::vector<PTorsionParams> PTorsionParser(const std::string& filename);
    static std::vector<BondParams> HBondParser(const std::string& filename);
    static std::vector<AngleParams> HAngleParser(const std::string& filename);
    static NonbondedParams NonBondedParser(const std::string& filename);
};

// Everything a state file carries: the OpenMM-style <State> (positions; when
// present velocities, box vectors, step count and time) and this engine's
// restart extension -- the exact box bounds and the neighbour-list state (the
// positions the current list was built from and whether a rebuild is due),
// without which a restarted mixed-precision run would not continue bit for bit.
struct StateData {
    std::vector<Coords3D> positions, velocities, buildPositions;
    bool hasVelocities = false, hasBoxVectors = false, hasBoxBounds = false, hasListState = false;
    Coords3D boxA, boxB, boxC;       // <PeriodicBoxVectors> A, B, C (nm)
    Coords3D lb, ub;                 // <NexusMDBox>: box bounds as the engine uses them
    int rebuildDue = 0;
    long long stepCount = 0;
    double time = 0.0;               // ps
};

class StateXMLParser {
   public:
    // <State><Positions><Position x y z> in nm (StateXMLParser.cpp:17-37)
    static std::vector<Coords3D> parseStateXML(const std::string& filename);
    // (A.x, B.y, C.z) of <PeriodicBoxVectors> (StateXMLParser.cpp:39-60; the
    // reference defines it and never calls it)
    static Coords3D extractBoxSize(const std::string& filename);
    // Every section of StateData in one pass over the file; throws
    // std::runtime_error when the file cannot be read or parsed.
    static StateData parseState(const std::string& filename);
    // A restartable state file: every section of StateData that is present,
    // doubles written with 17 significant digits (they read back exactly).
    static void writeState(const std::string& filename, const StateData& state);
};

struct PDBAtom {
    int atomNum = 0;
    std::string atomName, resName;
    char chainID = 'A';
    int resSeq = 0;
    Coords3D coords{0, 0, 0};
    double occupancy = 0, tempFactor = 0;
    std::string element, recordType;
};

// Report files, byte for byte as the reference writes them.
class Reporter {
   public:
    // "PVFReport" + filename; step 0 truncates, later steps append
    void TestPVFReport(const std::string& filename, const std::vector<Coords3D>& positions,
                       const std::vector<Coords3D>& velocities, const std::vector<Coords3D>& forces, int step);
    // "TotalEnergy" + filename; row 0 in fixed notation with 10 decimals, later
    // rows in the stream's default notation (a fresh stream per call)
    void TotalEnergyReport(const std::string& baseFilename, double totalKEnergy, double totalPEnergy,
                           double totalEnergy, int step);
    // REMARK + CRYST1 header; builds the template from the input PDB's ATOM lines
    void pdbOutputGeneratorPart1(const std::string& inputFilename, const std::string& outputFilename,
                                 std::vector<PDBAtom>& outputTemplate);
    // appends MODEL n, the atoms at `positions` (nm, written in Angstrom), TER and ENDMDL
    void pdbOutputGeneratorPart2(const std::string& outputFilename, const std::vector<PDBAtom>& outputTemplate,
                                 const std::vector<Coords3D>& positions, int ModelNum);
    static std::string formatPDBLine(const PDBAtom& atom);
    static PDBAtom parsePDBLine(const std::string& line);
    // ATOM lines of a PDB file as the output template; remembers its CRYST1 line
    void outputTemplateExtractor(const std::string& inputFilename, std::vector<PDBAtom>& outputTemplate);

    // date written into the REMARK line; empty = today (local time, YYYY-MM-DD)
    std::string date_override;

   private:
    std::string pbc_line_;
};

// The dispatcher.  CreateDispatcher("GPU") runs the FP64 path (the drop-in),
// "GPU_MIXED" the mixed path; any other choice prints the reference's warning
// and falls back to "GPU" (the reference falls back to its CPU engine,
// TaskDispatcher.cpp:6-15).
class TaskDispatcher {
   public:
    virtual ~TaskDispatcher() = default;
    virtual void Simulate(const std::string& systemFilename, const std::string& stateFilename,
                          const std::string& inputFilename, const std::string& outputFilename, double StepSize,
                          int TotalSteps, int interval) = 0;
    static std::shared_ptr<TaskDispatcher> CreateDispatcher(const std::string& processorChoice);

    void EnableHarmonicBondForce() { harmonicBondForceEnabled = true; }
    void EnableHarmonicAngleForce() { harmonicAngleForceEnabled = true; }
    void EnablePeriodicTorsionForce() { periodicTorsionForceEnabled = true; }
    void EnableNonbondedForce() { nonbondedForceEnabled = true; }

    // State I/O beyond the reference (each off by default, which keeps the
    // reference's behaviour: v(0) = 0, Initializer.cpp:21; the box from the
    // positions' bounding box, PeriodicBoundaryCondition.cpp:11-88, the state's
    // PeriodicBoxVectors unused, Engine.cpp:45-48).
    // v(0) from the state's <Velocities>
    void EnableStateVelocities() { useStateVelocities = true; }
    // the box from the state's <PeriodicBoxVectors> (rectangular; origin 0)
    void EnableStateBoxVectors() { useStateBoxVectors = true; }
    // continue a checkpoint: velocities, the exact box and the neighbour-list
    // state from the file, step numbering from its stepCount (reports append)
    void EnableRestart() { restart = useStateVelocities = useStateBoxVectors = true; }
    // write a restartable state file every `interval` steps (0: only at the end)
    // and after the last step
    void EnableCheckpoint(const std::string& filename, int interval) {
        checkpointFilename = filename;
        checkpointInterval = interval;
    }

    // Settings of this engine with no counterpart in the reference.
    int device = 0;                 // HIP device
    double skin = 0.1;              // Verlet buffer (nm); the pair set within the cutoff is unchanged by it
    bool echo_steps = true;         // print "Simulation Step #k is completed" as the reference does
    bool async_reports = true;      // report steps without waiting; files written by a writer thread
    bool useStateVelocities = false, useStateBoxVectors = false, restart = false;
    std::string checkpointFilename;
    int checkpointInterval = 0;

    // Wall-clock seconds of the last Simulate's ingest, stage by stage: the
    // system XML (one pass), the state XML, the topology (validation and the
    // 1-2/1-3/1-4 exclusions), box and state, and the device set-up (upload,
    // first neighbour-list build, PME tables).
    struct IngestTimes {
        double parse_system = 0, parse_state = 0, topology = 0, box_state = 0, device_setup = 0;
        long long atoms = 0;
    };
    IngestTimes lastIngest;

   protected:
    virtual void initializeParameters();
    std::vector<Coords3D> atomPositions;
    std::vector<double> masses;
    std::vector<PTorsionParams> torsionParams;
    std::vector<BondParams> bondParams;
    std::vector<AngleParams> angleParams;
    NonbondedParams nonbondedParams;
    bool harmonicBondForceEnabled = false;
    bool harmonicAngleForceEnabled = false;
    bool periodicTorsionForceEnabled = false;
    bool nonbondedForceEnabled = false;
};

class GpuTaskDispatcher : public TaskDispatcher {
   public:
    explicit GpuTaskDispatcher(int precision) : precision_(precision) {}
    void Simulate(const std::string& systemFilename, const std::string& stateFilename,
                  const std::string& inputFilename, const std::string& outputFilename, double StepSize,
                  int TotalSteps, int interval) override;
    int precision() const { return precision_; }

   private:
    int precision_;
};

}  // namespace nexusmd
