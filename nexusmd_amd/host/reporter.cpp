// reporter.cpp -- the reference's report files, byte for byte.
//
// Formats (Reporter.cpp:26-230 of the reference; kept identical because
// downstream tools read these files):
//   PVFReport<name>    "Step<k>", a header row, then per atom
//                      "<id>\t(x, y, z)\t(vx, vy, vz)\t(fx, fy, fz)" in fixed
//                      notation, 10 decimals, width 10; a blank line
//   TotalEnergy<name>  header and row 0 in fixed notation with 10 decimals;
//                      later rows in the default stream notation (each call
//                      opens a fresh stream); widths 5 / 25 / 30 / 25
//   <name>.pdb         REMARK and CRYST1 lines, then per report
//                      "MODEL        <n>", one line per ATOM of the input
//                      template with coordinates x10 (nm -> Angstrom), TER,
//                      ENDMDL
#include <chrono>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <set>
#include <sstream>

#include "nexusmd_host.h"

namespace nexusmd {

namespace {

const std::set<std::string>& amino_acids() {
    static const std::set<std::string> s = {"ALA", "ARG", "ASN", "ASP", "CYS", "GLN", "GLU", "GLY", "HIS", "ILE",
                                            "LEU", "LYS", "MET", "PHE", "PRO", "SER", "THR", "TRP", "TYR", "VAL"};
    return s;
}

// Element symbol: the first character of the atom name, plus the second when it is lower case.
std::string element_of(const std::string& name) {
    if (name.empty()) return "";
    std::string e(1, name[0]);
    if (name.size() > 1 && std::islower(static_cast<unsigned char>(name[1]))) e += name[1];
    return e;
}

std::string today() {
    const std::time_t t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    std::tm tm{};
    localtime_r(&t, &tm);
    char buf[16];
    std::strftime(buf, sizeof(buf), "%Y-%m-%d", &tm);
    return buf;
}

}  // namespace

void Reporter::TestPVFReport(const std::string& filename, const std::vector<Coords3D>& positions,
                             const std::vector<Coords3D>& velocities, const std::vector<Coords3D>& forces, int step) {
    const std::string name = "PVFReport" + filename;
    std::ofstream out(name, step == 0 ? std::ios::out : (std::ios::out | std::ios::app));
    if (!out.is_open()) {
        std::cerr << "Unable to open the file: " << name << std::endl;
        return;
    }
    out << "Step" << step << "\n";
    out << "ID\tPosition(nm)\t\t\tVelocity(nm / ps)\t\tForce(kJ / mol * nm)\n";
    out << std::fixed << std::setprecision(10);
    auto triple = [&](const Coords3D& v) {
        out << "(" << std::setw(10) << v[0] << ", " << std::setw(10) << v[1] << ", " << std::setw(10) << v[2] << ")";
    };
    for (size_t i = 0; i < positions.size(); ++i) {
        out << i + 1 << "\t";
        triple(positions[i]);
        out << "\t";
        triple(velocities[i]);
        out << "\t";
        triple(forces[i]);
        out << "\n";
    }
    out << "\n";
}

void Reporter::TotalEnergyReport(const std::string& baseFilename, double ke, double pe, double te, int step) {
    const std::string name = "TotalEnergy" + baseFilename;
    std::ofstream out(name, step == 0 ? std::ios::out : (std::ios::out | std::ios::app));
    if (!out.is_open()) {
        std::cerr << "Unable to open the file: " << name << std::endl;
        return;
    }
    if (step == 0) {
        out << std::fixed << std::setprecision(10);
        out << "Step\tTotal Kinetic Energy (kJ/mol)\tTotal Potential Energy (kJ/mol)\tTotal Energy (kJ/mol)\n";
    }
    out << std::setw(5) << step << std::setw(25) << ke << std::setw(30) << pe << std::setw(25) << te << "\n";
}

PDBAtom Reporter::parsePDBLine(const std::string& line) {
    PDBAtom a;
    std::istringstream ss(line);
    ss >> a.recordType >> a.atomNum >> a.atomName >> a.resName >> a.resSeq >> a.coords[0] >> a.coords[1] >>
        a.coords[2] >> a.occupancy >> a.tempFactor;
    a.chainID = 'A';
    a.element = element_of(a.atomName);
    a.recordType = amino_acids().count(a.resName) ? "ATOM  " : "HETATM";
    return a;
}

std::string Reporter::formatPDBLine(const PDBAtom& a) {
    std::ostringstream ss;
    ss << std::left << std::setw(6) << a.recordType << std::right << std::setw(5) << a.atomNum << " " << std::left
       << std::setw(4) << a.atomName << std::right << std::setw(1) << " " << std::setw(3) << a.resName << " "
       << std::setw(1) << a.chainID << std::setw(4) << std::right << a.resSeq << std::setw(1) << " "
       << "   " << std::fixed << std::setprecision(3) << std::right << std::setw(8) << a.coords[0] << std::right
       << std::setw(8) << a.coords[1] << std::right << std::setw(8) << a.coords[2] << std::fixed
       << std::setprecision(2) << std::setw(6) << a.occupancy << std::setw(6) << a.tempFactor << "          "
       << std::left << std::setw(2) << a.element;
    return ss.str();
}

void Reporter::outputTemplateExtractor(const std::string& inputFilename, std::vector<PDBAtom>& outputTemplate) {
    std::ifstream in(inputFilename);
    if (!in) {
        std::cerr << "Failed to open input file." << std::endl;
        return;
    }
    std::string line;
    while (std::getline(in, line)) {
        if (line.compare(0, 4, "ATOM") == 0) outputTemplate.push_back(parsePDBLine(line));
        else if (line.compare(0, 6, "CRYST1") == 0) pbc_line_ = line;
    }
}

void Reporter::pdbOutputGeneratorPart1(const std::string& inputFilename, const std::string& outputFilename,
                                       std::vector<PDBAtom>& outputTemplate) {
    std::ofstream out(outputFilename);
    outputTemplateExtractor(inputFilename, outputTemplate);
    out << "REMARK   Generated by Nexaus 1.0.0, " << (date_override.empty() ? today() : date_override) << std::endl;
    out << pbc_line_ << std::endl;
}

void Reporter::pdbOutputGeneratorPart2(const std::string& outputFilename, const std::vector<PDBAtom>& tmpl,
                                       const std::vector<Coords3D>& positions, int ModelNum) {
    std::ofstream out(outputFilename, std::ios::out | std::ios::app);
    if (!out.is_open()) {
        std::cerr << "Unable to open the file: " << outputFilename << std::endl;
        return;
    }
    std::cout << ModelNum << "   ";
    out << "MODEL        " << ModelNum << std::endl;
    for (size_t i = 0; i < tmpl.size(); ++i) {
        PDBAtom a = tmpl[i];
        const size_t k = static_cast<size_t>(a.atomNum - 1);
        if (a.atomNum < 1 || k >= positions.size())
            throw std::runtime_error("PDB atom number " + std::to_string(a.atomNum) + " has no position");
        a.coords = Coords3D{positions[k][0] * 10.0, positions[k][1] * 10.0, positions[k][2] * 10.0};
        out << formatPDBLine(a) << std::endl;
        if (i + 1 == tmpl.size()) {
            out << std::left << std::setw(6) << "TER" << std::right << std::setw(5) << a.atomNum + 1 << " "
                << std::left << std::setw(4) << "" << std::right << std::setw(1) << " " << std::setw(3) << a.resName
                << " " << std::setw(1) << a.chainID << std::setw(4) << std::right << a.resSeq << std::endl;
        }
    }
    out << "ENDMDL" << std::endl;
}

}  // namespace nexusmd
