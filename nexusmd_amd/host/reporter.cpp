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
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <ctime>
#include <future>
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

namespace {

// Appends v in the stream's fixed notation with `prec` decimals, right-aligned
// in `width` characters: std::to_chars in fixed notation with a precision is
// specified to give printf's "%.*f", which is what the stream writes in the
// "C" locale.  Non-finite values go through printf itself.
void append_fixed(std::string& buf, double v, int prec, int width) {
    char tmp[400];
    int len = 0;
    if (std::isfinite(v)) {
        const auto r = std::to_chars(tmp, tmp + sizeof(tmp), v, std::chars_format::fixed, prec);
        len = r.ec == std::errc() ? static_cast<int>(r.ptr - tmp) : -1;
    }
    if (len < 0 || !std::isfinite(v)) len = std::snprintf(tmp, sizeof(tmp), "%.*f", prec, v);
    if (len < width) buf.append(static_cast<size_t>(width - len), ' ');
    buf.append(tmp, static_cast<size_t>(len));
}

// The PVF rows of atoms [a0, a1): "<id>\t(x, y, z)\t(vx, vy, vz)\t(fx, fy, fz)\n",
// fixed notation, 10 decimals, width 10.
std::string pvf_rows(const std::vector<Coords3D>& P, const std::vector<Coords3D>& V, const std::vector<Coords3D>& F,
                     size_t a0, size_t a1) {
    std::string buf;
    buf.reserve((a1 - a0) * 150);
    auto triple = [&](const Coords3D& c) {
        buf += '(';
        append_fixed(buf, c[0], 10, 10);
        buf += ", ";
        append_fixed(buf, c[1], 10, 10);
        buf += ", ";
        append_fixed(buf, c[2], 10, 10);
        buf += ')';
    };
    char id[32];
    for (size_t i = a0; i < a1; ++i) {
        const auto r = std::to_chars(id, id + sizeof(id), i + 1);
        buf.append(id, static_cast<size_t>(r.ptr - id));
        buf += '\t';
        triple(P[i]);
        buf += '\t';
        triple(V[i]);
        buf += '\t';
        triple(F[i]);
        buf += '\n';
    }
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
    if (velocities.size() < positions.size() || forces.size() < positions.size())
        throw std::invalid_argument("TestPVFReport: fewer velocities or forces than positions");
    std::string head = "Step" + std::to_string(step) + "\n";
    head += "ID\tPosition(nm)\t\t\tVelocity(nm / ps)\t\tForce(kJ / mol * nm)\n";
    out.write(head.data(), static_cast<std::streamsize>(head.size()));
    // rows formatted in parallel blocks for large systems, written in order
    const size_t n = positions.size();
    const size_t nblk = n >= 16384 ? std::min<size_t>(8, n / 4096) : 1;
    std::vector<std::future<std::string>> parts;
    for (size_t b = 1; b < nblk; ++b)
        parts.push_back(std::async(std::launch::async, pvf_rows, std::cref(positions), std::cref(velocities),
                                   std::cref(forces), n * b / nblk, n * (b + 1) / nblk));
    const std::string first = pvf_rows(positions, velocities, forces, 0, n / nblk);
    out.write(first.data(), static_cast<std::streamsize>(first.size()));
    for (auto& f : parts) {
        const std::string rows = f.get();
        out.write(rows.data(), static_cast<std::streamsize>(rows.size()));
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
    // the reference's manipulator chain as one C format: left 6 / right 5 / left 4 /
    // right 3 / char / right 4, then fixed 8.3 x3 and 6.2 x2, left 2
    char line[256];
    const int len = std::snprintf(line, sizeof(line), "%-6s%5d %-4s %3s %c%4d    %8.3f%8.3f%8.3f%6.2f%6.2f          %-2s",
                                  a.recordType.c_str(), a.atomNum, a.atomName.c_str(), a.resName.c_str(), a.chainID,
                                  a.resSeq, a.coords[0], a.coords[1], a.coords[2], a.occupancy, a.tempFactor,
                                  a.element.c_str());
    if (len > 0 && static_cast<size_t>(len) < sizeof(line)) return std::string(line, static_cast<size_t>(len));
    std::ostringstream ss;     // very long fields: the stream chain itself
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
    // the model is formatted into one buffer and written at once (the same bytes
    // as a flush after every line)
    std::string buf = "MODEL        " + std::to_string(ModelNum) + "\n";
    buf.reserve(buf.size() + (tmpl.size() + 2) * 81);
    for (size_t i = 0; i < tmpl.size(); ++i) {
        PDBAtom a = tmpl[i];
        const size_t k = static_cast<size_t>(a.atomNum - 1);
        if (a.atomNum < 1 || k >= positions.size())
            throw std::runtime_error("PDB atom number " + std::to_string(a.atomNum) + " has no position");
        a.coords = Coords3D{positions[k][0] * 10.0, positions[k][1] * 10.0, positions[k][2] * 10.0};
        buf += formatPDBLine(a);
        buf += "\n";
        if (i + 1 == tmpl.size()) {
            char ter[128];
            const int len = std::snprintf(ter, sizeof(ter), "%-6s%5d %-4s %3s %c%4d\n", "TER", a.atomNum + 1, "",
                                          a.resName.c_str(), a.chainID, a.resSeq);
            if (len > 0 && static_cast<size_t>(len) < sizeof(ter)) {
                buf.append(ter, static_cast<size_t>(len));
            } else {
                std::ostringstream ss;
                ss << std::left << std::setw(6) << "TER" << std::right << std::setw(5) << a.atomNum + 1 << " "
                   << std::left << std::setw(4) << "" << std::right << std::setw(1) << " " << std::setw(3)
                   << a.resName << " " << std::setw(1) << a.chainID << std::setw(4) << std::right << a.resSeq << "\n";
                buf += ss.str();
            }
        }
    }
    buf += "ENDMDL\n";
    out.write(buf.data(), static_cast<std::streamsize>(buf.size()));
}

}  // namespace nexusmd
