// pymodule.cpp -- the Python module `NexusMD`, exposing the same names as the
// reference's SWIG wrapper (wrapper/NexusMD.i: TaskDispatcher with its factory,
// the Enable*Force calls and Simulate; Coords3D; the SystemXMLParser statics),
// so that scripts written for the reference (test/NexusMDTest.py) run
// unchanged.  C++ exceptions surface as Python RuntimeError / ValueError.
#include <pybind11/pybind11.h>
#include <pybind11/numpy.h>
#include <pybind11/stl.h>

#include "nexusmd_host.h"

namespace py = pybind11;
using namespace nexusmd;

PYBIND11_MODULE(NexusMD, m) {
    m.doc() = "NexusMD force-and-integrate engine on AMD Instinct MI355X (HIP, gfx950)";

    py::class_<TaskDispatcher, std::shared_ptr<TaskDispatcher>>(m, "TaskDispatcher")
        .def_static("CreateDispatcher", &TaskDispatcher::CreateDispatcher, py::arg("processorChoice"))
        .def("EnableHarmonicBondForce", &TaskDispatcher::EnableHarmonicBondForce)
        .def("EnableHarmonicAngleForce", &TaskDispatcher::EnableHarmonicAngleForce)
        .def("EnablePeriodicTorsionForce", &TaskDispatcher::EnablePeriodicTorsionForce)
        .def("EnableNonbondedForce", &TaskDispatcher::EnableNonbondedForce)
        .def("Simulate", &TaskDispatcher::Simulate, py::arg("systemFilename"), py::arg("stateFilename"),
             py::arg("inputFilename"), py::arg("outputFilename"), py::arg("StepSize"), py::arg("TotalSteps"),
             py::arg("interval"), py::call_guard<py::gil_scoped_release>())
        .def("EnableStateVelocities", &TaskDispatcher::EnableStateVelocities)
        .def("EnableStateBoxVectors", &TaskDispatcher::EnableStateBoxVectors)
        .def("EnableRestart", &TaskDispatcher::EnableRestart)
        .def("EnableCheckpoint", &TaskDispatcher::EnableCheckpoint, py::arg("filename"), py::arg("interval") = 0)
        .def_readwrite("device", &TaskDispatcher::device)
        .def_readwrite("skin", &TaskDispatcher::skin)
        .def_readwrite("echo_steps", &TaskDispatcher::echo_steps)
        .def_readwrite("async_reports", &TaskDispatcher::async_reports)
        .def("IngestTimes", [](const TaskDispatcher& d) {
            const auto& t = d.lastIngest;
            py::dict r;
            r["atoms"] = t.atoms;
            r["parse_system_s"] = t.parse_system;
            r["parse_state_s"] = t.parse_state;
            r["topology_s"] = t.topology;
            r["box_state_s"] = t.box_state;
            r["device_setup_s"] = t.device_setup;
            return r;
        });

    // Coords3D with the reference's operations (SWIG exports it, NexusMD.i:18)
    py::class_<Coords3D>(m, "Coords3D")
        .def(py::init<>())
        .def(py::init<double, double, double>(), py::arg("x"), py::arg("y"), py::arg("z"))
        // from any sequence of three numbers (a list, tuple or numpy row), so that
        // lists of positions pass where the C++ side takes std::vector<Coords3D>
        .def(py::init([](const py::sequence& s) {
            if (py::len(s) != 3) throw py::value_error("Coords3D needs three components");
            return Coords3D(s[0].cast<double>(), s[1].cast<double>(), s[2].cast<double>());
        }))
        .def("__getitem__",
             [](const Coords3D& c, int i) {
                 if (i < 0) i += 3;
                 if (i < 0 || i > 2) throw py::index_error("Coords3D index out of range");
                 return c[i];
             })
        .def("__setitem__",
             [](Coords3D& c, int i, double v) {
                 if (i < 0) i += 3;
                 if (i < 0 || i > 2) throw py::index_error("Coords3D index out of range");
                 c[i] = v;
             })
        .def("__len__", [](const Coords3D&) { return 3; })
        .def("__eq__", [](const Coords3D& a, const Coords3D& b) { return a == b; })
        .def("__ne__", [](const Coords3D& a, const Coords3D& b) { return a != b; })
        .def("__pos__", [](const Coords3D& a) { return +a; })
        .def("__neg__", [](const Coords3D& a) { return -a; })
        .def("__add__", [](const Coords3D& a, const Coords3D& b) { return a + b; })
        .def("__sub__", [](const Coords3D& a, const Coords3D& b) { return a - b; })
        .def("__iadd__", [](Coords3D& a, const Coords3D& b) { return a += b; })
        .def("__isub__", [](Coords3D& a, const Coords3D& b) { return a -= b; })
        .def("__mul__", [](const Coords3D& a, double s) { return a * s; })
        .def("__imul__", [](Coords3D& a, double s) { return a *= s; })
        .def("__truediv__", [](const Coords3D& a, double s) { return a / s; })
        .def("__itruediv__", [](Coords3D& a, double s) { return a /= s; })
        .def("dot", &Coords3D::dot)
        .def("cross", &Coords3D::cross)
        .def("length", &Coords3D::length)
        .def("normalize", &Coords3D::normalize)
        .def("__repr__", [](const Coords3D& c) {
            return "Coords3D(" + std::to_string(c[0]) + ", " + std::to_string(c[1]) + ", " + std::to_string(c[2]) + ")";
        });
    py::implicitly_convertible<py::list, Coords3D>();
    py::implicitly_convertible<py::tuple, Coords3D>();
    py::implicitly_convertible<py::array, Coords3D>();

    py::class_<BondParams>(m, "BondParams")
        .def_readonly("p1", &BondParams::p1).def_readonly("p2", &BondParams::p2)
        .def_readonly("d", &BondParams::d).def_readonly("k", &BondParams::k);
    py::class_<AngleParams>(m, "AngleParams")
        .def_readonly("p1", &AngleParams::p1).def_readonly("p2", &AngleParams::p2)
        .def_readonly("p3", &AngleParams::p3).def_readonly("a", &AngleParams::a)
        .def_readonly("k", &AngleParams::k);
    py::class_<PTorsionParams>(m, "PTorsionParams")
        .def_readonly("p1", &PTorsionParams::p1).def_readonly("p2", &PTorsionParams::p2)
        .def_readonly("p3", &PTorsionParams::p3).def_readonly("p4", &PTorsionParams::p4)
        .def_readonly("k", &PTorsionParams::k).def_readonly("phase", &PTorsionParams::phase)
        .def_readonly("periodicity", &PTorsionParams::periodicity);
    py::class_<NonbondedParticle>(m, "NonbondedParticle")
        .def_readonly("q", &NonbondedParticle::q).def_readonly("sig", &NonbondedParticle::sig)
        .def_readonly("eps", &NonbondedParticle::eps);
    py::class_<NonbondedException>(m, "NonbondedException")
        .def_readonly("p1", &NonbondedException::p1).def_readonly("p2", &NonbondedException::p2)
        .def_readonly("q", &NonbondedException::q).def_readonly("sig", &NonbondedException::sig)
        .def_readonly("eps", &NonbondedException::eps);
    py::class_<NonbondedParams>(m, "NonbondedParams")
        .def_readonly("alpha", &NonbondedParams::alpha)
        .def_readonly("cutoff", &NonbondedParams::cutoff)
        .def_readonly("ewaldTolerance", &NonbondedParams::ewaldTolerance)
        .def_readonly("rfDielectric", &NonbondedParams::rfDielectric)
        .def_readonly("switchingDistance", &NonbondedParams::switchingDistance)
        .def_readonly("useSwitchingFunction", &NonbondedParams::useSwitchingFunction)
        .def_readonly("method", &NonbondedParams::method)
        .def_readonly("nx", &NonbondedParams::nx).def_readonly("ny", &NonbondedParams::ny)
        .def_readonly("nz", &NonbondedParams::nz)
        .def_readonly("particles", &NonbondedParams::particles)
        .def_readonly("exceptions", &NonbondedParams::exceptions);

    py::class_<SystemXMLParser>(m, "SystemXMLParser")
        .def_static("MassesParser", &SystemXMLParser::MassesParser)
        .def_static("PTorsionParser", &SystemXMLParser::PTorsionParser)
        .def_static("HBondParser", &SystemXMLParser::HBondParser)
        .def_static("HAngleParser", &SystemXMLParser::HAngleParser)
        .def_static("NonBondedParser", &SystemXMLParser::NonBondedParser);
    py::class_<StateData>(m, "StateData")
        .def(py::init<>())
        .def_readwrite("positions", &StateData::positions)
        .def_readwrite("velocities", &StateData::velocities)
        .def_readwrite("buildPositions", &StateData::buildPositions)
        .def_readwrite("hasVelocities", &StateData::hasVelocities)
        .def_readwrite("hasBoxVectors", &StateData::hasBoxVectors)
        .def_readwrite("hasBoxBounds", &StateData::hasBoxBounds)
        .def_readwrite("hasListState", &StateData::hasListState)
        .def_readwrite("boxA", &StateData::boxA)
        .def_readwrite("boxB", &StateData::boxB)
        .def_readwrite("boxC", &StateData::boxC)
        .def_readwrite("lb", &StateData::lb)
        .def_readwrite("ub", &StateData::ub)
        .def_readwrite("rebuildDue", &StateData::rebuildDue)
        .def_readwrite("stepCount", &StateData::stepCount)
        .def_readwrite("time", &StateData::time);
    py::class_<StateXMLParser>(m, "StateXMLParser")
        .def_static("parseStateXML", &StateXMLParser::parseStateXML)
        .def_static("extractBoxSize", &StateXMLParser::extractBoxSize)
        .def_static("parseState", &StateXMLParser::parseState)
        .def_static("writeState", &StateXMLParser::writeState);

    py::class_<Reporter>(m, "Reporter")
        .def(py::init<>())
        .def_readwrite("date_override", &Reporter::date_override)
        .def("TestPVFReport", &Reporter::TestPVFReport)
        .def("TotalEnergyReport", &Reporter::TotalEnergyReport)
        .def("pdbOutputGeneratorPart1",
             [](Reporter& r, const std::string& in, const std::string& out) {
                 std::vector<PDBAtom> t;
                 r.pdbOutputGeneratorPart1(in, out, t);
                 return t.size();
             })
        .def("pdbReport",
             [](Reporter& r, const std::string& in, const std::string& out, const std::vector<Coords3D>& pos,
                int model, bool first) {
                 // Part1 (when first), then Part2 with the template of `in`
                 std::vector<PDBAtom> t;
                 if (first) r.pdbOutputGeneratorPart1(in, out, t);
                 else r.outputTemplateExtractor(in, t);
                 r.pdbOutputGeneratorPart2(out, t, pos, model);
             });
}
