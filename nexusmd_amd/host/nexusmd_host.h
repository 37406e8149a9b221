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
#include <memory>
#include <string>
#include <vector>

#include "nexusmd.h"

namespace nexusmd {

using Coords3D = std::array<double, 3>;

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

class StateXMLParser {
   public:
    // <State><Positions><Position x y z> in nm (StateXMLParser.cpp:17-37)
    static std::vector<Coords3D> parseStateXML(const std::string& filename);
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

    // Settings of this engine with no counterpart in the reference.
    int device = 0;                 // HIP device
    double skin = 0.1;              // Verlet buffer (nm); the pair set within the cutoff is unchanged by it
    bool echo_steps = true;         // print "Simulation Step #k is completed" as the reference does

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
