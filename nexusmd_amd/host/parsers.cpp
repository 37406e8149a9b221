// parsers.cpp -- system and state XML readers (one pass per file).
//
// Element selection follows the reference's parsers: the first <Particles>
// of <System> gives the masses (SystemXMLParser.cpp:18-35); every <Force>
// whose type is HarmonicBondForce / HarmonicAngleForce / PeriodicTorsionForce
// contributes the children of its first <Bonds> / <Angles> / <Torsions>
// (:37-130); only the first NonbondedForce is read, attributes, <Particles>
// and <Exceptions> (:133-196); positions come from the first <Positions> of
// <State> (StateXMLParser.cpp:17-37).  Missing attributes read as 0.
#include <cstdio>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>

#include "nexusmd_host.h"
#include "xml_reader.h"

namespace nexusmd {

namespace {

// Tracks, for every open element, how many times each child name has been seen,
// so that "first <X> under this element" can be decided while streaming.
class Occurrence {
   public:
    // Returns the 1-based occurrence of the element that just started at depth d.
    int enter(const std::vector<std::string>& path) {
        const size_t d = path.size();
        if (counts_.size() < d) counts_.resize(d);
        if (first_.size() < d) first_.resize(d);
        counts_[d - 1].clear();                       // fresh child counts for this element
        int occ = 1;
        if (d >= 2) occ = ++counts_[d - 2][path[d - 1]];
        else occ = ++roots_[path[0]];
        first_[d - 1] = occ;
        return occ;
    }
    int at(size_t depth) const { return first_[depth]; }   // occurrence of the element at depth

   private:
    std::vector<std::map<std::string, int>> counts_;
    std::vector<int> first_;
    std::map<std::string, int> roots_;
};

bool is_path(const std::vector<std::string>& p, std::initializer_list<const char*> names) {
    if (p.size() != names.size()) return false;
    size_t i = 0;
    for (const char* n : names)
        if (p[i++] != n) return false;
    return true;
}

}  // namespace

SystemDescription SystemXMLParser::Parse(const std::string& filename) {
    SystemDescription out;
    Occurrence occ;
    std::string force_type;
    bool first_nb = false;      // inside the first NonbondedForce
    bool seen_nb = false;
    std::string err;
    auto on_start = [&](const std::vector<std::string>& p, const XmlAttrs& a) {
        const int o = occ.enter(p);
        (void)o;
        if (p.empty() || p[0] != "System" || occ.at(0) != 1) return;
        if (is_path(p, {"System", "Particles", "Particle"})) {
            if (occ.at(1) == 1) out.masses.push_back(a.as_double("mass"));
            return;
        }
        if (p.size() < 3 || p[1] != "Forces" || occ.at(1) != 1 || p[2] != "Force") return;
        if (p.size() == 3) {
            force_type = a.as_string("type");
            first_nb = false;
            if (force_type == "NonbondedForce" && !seen_nb) {
                seen_nb = first_nb = true;
                NonbondedParams& nb = out.nonbonded;
                nb.present = true;
                nb.alpha = a.as_double("alpha");
                nb.cutoff = a.as_double("cutoff");
                nb.dispersionCorrection = a.as_int("dispersionCorrection");
                nb.ewaldTolerance = a.as_double("ewaldTolerance");
                nb.exceptionsUsePeriodic = a.as_int("exceptionsUsePeriodic");
                nb.forceGroup = a.as_int("forceGroup");
                nb.includeDirectSpace = a.as_int("includeDirectSpace");
                nb.ljAlpha = a.as_int("ljAlpha");
                nb.ljnx = a.as_int("ljnx");
                nb.ljny = a.as_int("ljny");
                nb.ljnz = a.as_int("ljnz");
                nb.method = a.as_int("method");
                nb.nx = a.as_int("nx");
                nb.ny = a.as_int("ny");
                nb.nz = a.as_int("nz");
                nb.recipForceGroup = a.as_int("recipForceGroup");
                nb.rfDielectric = a.as_double("rfDielectric");
                nb.switchingDistance = a.as_double("switchingDistance");
                nb.useSwitchingFunction = a.as_int("useSwitchingFunction");
                nb.version = a.as_int("version");
            }
            return;
        }
        if (p.size() != 5 || occ.at(3) != 1) return;
        const std::string& box = p[3];
        const std::string& item = p[4];
        if (force_type == "HarmonicBondForce" && box == "Bonds" && item == "Bond") {
            out.bonds.push_back(nx_bond{a.as_int("p1"), a.as_int("p2"), a.as_double("d"), a.as_double("k")});
        } else if (force_type == "HarmonicAngleForce" && box == "Angles" && item == "Angle") {
            out.angles.push_back(
                nx_angle{a.as_int("p1"), a.as_int("p2"), a.as_int("p3"), a.as_double("a"), a.as_double("k")});
        } else if (force_type == "PeriodicTorsionForce" && box == "Torsions" && item == "Torsion") {
            out.torsions.push_back(nx_torsion{a.as_int("p1"), a.as_int("p2"), a.as_int("p3"), a.as_int("p4"),
                                              a.as_double("k"), a.as_double("phase"), a.as_int("periodicity")});
        } else if (first_nb && box == "Particles" && item == "Particle") {
            out.nonbonded.particles.push_back(
                NonbondedParticle{a.as_double("q"), a.as_double("sig"), a.as_double("eps")});
        } else if (first_nb && box == "Exceptions" && item == "Exception") {
            out.nonbonded.exceptions.push_back(NonbondedException{a.as_int("p1"), a.as_int("p2"), a.as_double("q"),
                                                                  a.as_double("sig"), a.as_double("eps")});
        }
    };
    if (!xml_scan_file(filename, on_start, &err)) throw std::runtime_error("XML [" + filename + "]: " + err);
    return out;
}

namespace {
template <class F>
auto or_empty(const std::string& filename, F&& f) -> decltype(f(SystemDescription{})) {
    try {
        return f(SystemXMLParser::Parse(filename));
    } catch (const std::exception& e) {
        std::cerr << "XML [" << filename << "] parsed with errors, Error description: " << e.what() << "\n";
        return decltype(f(SystemDescription{})){};
    }
}
}  // namespace

std::vector<double> SystemXMLParser::MassesParser(const std::string& f) {
    return or_empty(f, [](SystemDescription d) { return d.masses; });
}
std::vector<PTorsionParams> SystemXMLParser::PTorsionParser(const std::string& f) {
    return or_empty(f, [](SystemDescription d) { return d.torsions; });
}
std::vector<BondParams> SystemXMLParser::HBondParser(const std::string& f) {
    return or_empty(f, [](SystemDescription d) { return d.bonds; });
}
std::vector<AngleParams> SystemXMLParser::HAngleParser(const std::string& f) {
    return or_empty(f, [](SystemDescription d) { return d.angles; });
}
NonbondedParams SystemXMLParser::NonBondedParser(const std::string& f) {
    return or_empty(f, [](SystemDescription d) { return d.nonbonded; });
}

std::vector<Coords3D> StateXMLParser::parseStateXML(const std::string& filename) {
    std::vector<Coords3D> pos;
    Occurrence occ;
    std::string err;
    auto on_start = [&](const std::vector<std::string>& p, const XmlAttrs& a) {
        occ.enter(p);
        if (is_path(p, {"State", "Positions", "Position"}) && occ.at(0) == 1 && occ.at(1) == 1)
            pos.push_back(Coords3D{a.as_double("x"), a.as_double("y"), a.as_double("z")});
    };
    try {
        if (!xml_scan_file(filename, on_start, &err)) throw std::runtime_error(err);
    } catch (const std::exception& e) {
        std::cerr << "XML [" << filename << "] parsed with errors, Error description: " << e.what() << "\n";
        return {};
    }
    return pos;
}

Coords3D StateXMLParser::extractBoxSize(const std::string& filename) {
    try {
        const StateData s = parseState(filename);
        if (!s.hasBoxVectors) std::cerr << "PeriodicBoxVectors not found in XML [" << filename << "]\n";
        return Coords3D(s.boxA[0], s.boxB[1], s.boxC[2]);
    } catch (const std::exception& e) {
        std::cerr << "XML [" << filename << "] parsed with errors, Error description: " << e.what() << "\n";
        return Coords3D();
    }
}

StateData StateXMLParser::parseState(const std::string& filename) {
    StateData s;
    Occurrence occ;
    std::string err;
    auto xyz = [](const XmlAttrs& a) { return Coords3D(a.as_double("x"), a.as_double("y"), a.as_double("z")); };
    auto on_start = [&](const std::vector<std::string>& p, const XmlAttrs& a) {
        occ.enter(p);
        if (p.empty() || p[0] != "State" || occ.at(0) != 1) return;
        if (p.size() == 1) {
            s.stepCount = (long long)a.as_double("stepCount");
            s.time = a.as_double("time");
            return;
        }
        if (occ.at(1) != 1) return;       // the first section of each name
        const std::string& sec = p[1];
        if (p.size() == 2) {
            if (sec == "NexusMDBox") {
                s.hasBoxBounds = true;
                s.lb = Coords3D(a.as_double("lbx"), a.as_double("lby"), a.as_double("lbz"));
                s.ub = Coords3D(a.as_double("ubx"), a.as_double("uby"), a.as_double("ubz"));
            } else if (sec == "NexusMDNeighborList") {
                s.hasListState = true;
                s.rebuildDue = a.as_int("rebuildDue");
            } else if (sec == "Velocities") {
                s.hasVelocities = true;
            } else if (sec == "PeriodicBoxVectors") {
                s.hasBoxVectors = true;
            }
            return;
        }
        if (p.size() != 3) return;
        const std::string& item = p[2];
        if (sec == "Positions" && item == "Position") s.positions.push_back(xyz(a));
        else if (sec == "Velocities" && item == "Velocity") s.velocities.push_back(xyz(a));
        else if (sec == "NexusMDNeighborList" && item == "BuildPosition") s.buildPositions.push_back(xyz(a));
        else if (sec == "PeriodicBoxVectors" && item == "A") s.boxA = xyz(a);
        else if (sec == "PeriodicBoxVectors" && item == "B") s.boxB = xyz(a);
        else if (sec == "PeriodicBoxVectors" && item == "C") s.boxC = xyz(a);
    };
    if (!xml_scan_file(filename, on_start, &err)) throw std::runtime_error("XML [" + filename + "]: " + err);
    return s;
}

void StateXMLParser::writeState(const std::string& filename, const StateData& s) {
    FILE* f = std::fopen(filename.c_str(), "w");
    if (!f) throw std::runtime_error("cannot write " + filename);
    std::unique_ptr<FILE, int (*)(FILE*)> hold(f, std::fclose);
    auto row = [&](const char* tag, const Coords3D& v) {
        std::fprintf(f, "\t\t<%s x=\"%.17g\" y=\"%.17g\" z=\"%.17g\"/>\n", tag, v[0], v[1], v[2]);
    };
    std::fprintf(f, "<?xml version=\"1.0\" ?>\n<State stepCount=\"%lld\" time=\"%.17g\" type=\"State\" version=\"1\">\n",
                 s.stepCount, s.time);
    if (s.hasBoxVectors) {
        std::fprintf(f, "\t<PeriodicBoxVectors>\n");
        row("A", s.boxA);
        row("B", s.boxB);
        row("C", s.boxC);
        std::fprintf(f, "\t</PeriodicBoxVectors>\n");
    }
    if (s.hasBoxBounds)
        std::fprintf(f, "\t<NexusMDBox lbx=\"%.17g\" lby=\"%.17g\" lbz=\"%.17g\" ubx=\"%.17g\" uby=\"%.17g\" ubz=\"%.17g\"/>\n",
                     s.lb[0], s.lb[1], s.lb[2], s.ub[0], s.ub[1], s.ub[2]);
    std::fprintf(f, "\t<Positions>\n");
    for (const auto& v : s.positions) row("Position", v);
    std::fprintf(f, "\t</Positions>\n");
    if (s.hasVelocities) {
        std::fprintf(f, "\t<Velocities>\n");
        for (const auto& v : s.velocities) row("Velocity", v);
        std::fprintf(f, "\t</Velocities>\n");
    }
    if (s.hasListState) {
        std::fprintf(f, "\t<NexusMDNeighborList rebuildDue=\"%d\">\n", s.rebuildDue);
        for (const auto& v : s.buildPositions) row("BuildPosition", v);
        std::fprintf(f, "\t</NexusMDNeighborList>\n");
    }
    std::fprintf(f, "</State>\n");
    if (std::ferror(f)) throw std::runtime_error("write error on " + filename);
}

}  // namespace nexusmd
