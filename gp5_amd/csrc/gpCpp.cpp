// pybind11 module gpCpp: the reference's Python-facing classes and methods
// (src/gpCpp.cpp:16-370) over the C++ facade gp5pp.hpp, i.e. over the HIP
// kernels of libgp5.so.  Same class names, constructors and method names;
// differences:
//   * one GPU: processorGrid must be [1,1,1] (fftOrdering is accepted);
//   * an unknown stepper name raises (the reference dereferences null);
//   * the GIL is released while kernels run;
//   * field.save/load write the HDF5 field file serially (one writer);
//   * additions, all with reference-preserving defaults: GP and UnlockedLHY
//     models, model.energy, field.maxDensity, timeStepper.advanceLog.
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "gp5pp.hpp"

namespace py = pybind11;
using gp::complex_t;
using gp::real_t;

namespace pyInterface {

class geometry {
public:
    geometry(std::array<real_t, 3> left, std::array<real_t, 3> right, std::array<int, 3> shape)
        : _left(left), _right(right), _shape(shape)
    {
    }
    const std::array<real_t, 3> &getLeft() const { return _left; }
    const std::array<real_t, 3> &getRight() const { return _right; }
    const std::array<int, 3> &getShape() const { return _shape; }

private:
    std::array<real_t, 3> _left, _right;
    std::array<int, 3> _shape;
};

// src/gpCpp.cpp:41-149: owns the state tensor and the transform
class field {
public:
    field(const geometry &geo, int nComponents, std::array<int, 3> processorGrid, std::array<int, 3> fftOrdering)
        : _nComponents(nComponents)
    {
        gp::fourierTransformCreator fftC;
        fftC.setNComponents(nComponents);
        fftC.setDomain(std::make_shared<gp::domain>(geo.getLeft(), geo.getRight()));
        fftC.setGlobalMesh(std::make_shared<gp::mesh>(geo.getShape()));
        fftC.setProcessorGrid(processorGrid);
        fftC.setOrdering(fftOrdering);
        _fftOp = fftC.create();
        _state = std::make_shared<gp::tensor_t>(getDiscretization(), nComponents);
        _state->setZero();
    }
    std::shared_ptr<gp::discretization> getDiscretization() { return _fftOp->getDiscretizationRealSpace(); }
    int nComponents() const { return _nComponents; }
    std::array<int, 3> getLocalShape() { return getDiscretization()->getLocalMesh()->shape(); }
    std::array<int, 3> getOffset() { return {0, 0, 0}; }

    // F-ordered copy, as pyTools.h:93-103
    py::array_t<complex_t> toPyArray()
    {
        auto d = _state->dimensions();
        py::array_t<complex_t> a({(py::ssize_t)d[0], (py::ssize_t)d[1], (py::ssize_t)d[2], (py::ssize_t)d[3]},
                                 {(py::ssize_t)sizeof(complex_t), (py::ssize_t)(sizeof(complex_t) * d[0]),
                                  (py::ssize_t)(sizeof(complex_t) * d[0] * d[1]),
                                  (py::ssize_t)(sizeof(complex_t) * d[0] * d[1] * d[2])});
        _state->toHost(a.mutable_data());
        return a;
    }

    // any rank-4 complex array of the field's shape (pyTools.h:12-63)
    void setTensor(py::array_t<complex_t, py::array::f_style | py::array::forcecast> arr)
    {
        py::buffer_info info = arr.request();
        if (info.ndim != 4) throw std::runtime_error("Unexpected rank of the python array.");
        auto d = _state->dimensions();
        for (int i = 0; i < 4; ++i)
            if (info.shape[i] != d[i])
                throw std::runtime_error("array shape does not match the field (" + std::to_string(d[0]) + "," +
                                         std::to_string(d[1]) + "," + std::to_string(d[2]) + "," +
                                         std::to_string(d[3]) + ")");
        _state->fromHost(static_cast<const complex_t *>(info.ptr));
    }

    real_t getNorm(int c) { return gp::norm(*_state, c); }
    void normalize(int component, real_t N) { gp::normalize(N, *_state, component); }
    real_t maxDensity(int c)
    {
        double out = 0.0;
        gp::check(gp5_max_density(_state->ctx(), _state->buf(), c, &out));
        return out;
    }
    // src/gpCpp.cpp:105-113: the reference's HDF5 field file, written serially
    void save(const std::string &filename) { gp::save(*_state, filename, *getDiscretization()); }
    void load(const std::string &filename)
    {
        auto host = gp::loadHost(filename, *getDiscretization(), _nComponents);
        _state->fromHost(host.data());
    }

    std::shared_ptr<gp::tensor_t> getState() { return _state; }
    std::shared_ptr<gp::fourierTransform> getFFTOp() { return _fftOp; }

private:
    int _nComponents;
    std::shared_ptr<gp::fourierTransform> _fftOp;
    std::shared_ptr<gp::tensor_t> _state;
};

// src/gpCpp.cpp:151-172
class model {
public:
    virtual ~model() = default;
    void initFunctional(std::shared_ptr<gp::functional> func, field &psi)
    {
        func->setLaplacianOperator(std::make_shared<gp::operators::laplacian>(psi.getFFTOp()));
        func->setDiscretization(psi.getDiscretization());
        func->setNComponents(psi.nComponents());
        func->init();
    }
    virtual std::shared_ptr<gp::functional> getFunctional() = 0;
    void apply(field &source, field &dest, real_t time)
    {
        if (source.getState().get() == dest.getState().get())
            throw std::runtime_error("model.apply needs distinct source and destination fields");
        py::gil_scoped_release nogil;
        getFunctional()->apply(*source.getState(), *dest.getState(), time);
    }
    real_t energy(field &psi)
    {
        py::gil_scoped_release nogil;
        return getFunctional()->energy(*psi.getState());
    }
    int nComponents() { return getFunctional()->getNComponents(); }
};

// src/gpCpp.cpp:178-199 (coefficients -3, 5/2 unless given)
class LHY : public model {
public:
    explicit LHY(field &psi, real_t a = -3.0, real_t b = 2.5)
    {
        _func = std::make_shared<gp::LHYDropletFunctional>();
        _func->setCoefficients(a, b);
        initFunctional(_func, psi);
    }
    std::shared_ptr<gp::functional> getFunctional() override { return _func; }

private:
    std::shared_ptr<gp::LHYDropletFunctional> _func;
};

// gpFunctional with optional couplings, masses and harmonic trap (addition)
class GP : public model {
public:
    GP(field &psi, std::vector<std::vector<real_t>> couplings, std::vector<real_t> masses,
       std::vector<std::array<real_t, 3>> omegas)
    {
        _func = std::make_shared<gp::gpFunctional>();
        if (!couplings.empty()) _func->setCouplings(couplings);
        if (!masses.empty()) _func->setMasses(masses);
        if (!omegas.empty())
            _func->setExternalPotential(gp::harmonicPotential(omegas).create(psi.getDiscretization(), psi.nComponents()));
        initFunctional(_func, psi);
    }
    std::shared_ptr<gp::functional> getFunctional() override { return _func; }

private:
    std::shared_ptr<gp::gpFunctional> _func;
};

// LHYDropletUnlockedFunctional (addition)
class UnlockedLHY : public model {
public:
    UnlockedLHY(field &psi, real_t alpha, real_t beta, real_t eta)
    {
        if (psi.nComponents() != 2) throw std::runtime_error("Incompatible number of components");
        _func = std::make_shared<gp::LHYDropletUnlockedFunctional>(alpha, beta, eta);
        initFunctional(_func, psi);
    }
    std::shared_ptr<gp::functional> getFunctional() override { return _func; }

private:
    std::shared_ptr<gp::LHYDropletUnlockedFunctional> _func;
};

// src/gpCpp.cpp:202-287
class timeStepper {
public:
    timeStepper(real_t timeStep, std::string stepperName, bool isImaginary) : _timeStep(timeStep)
    {
        if (stepperName == "eulero") _stepper = std::make_shared<gp::euleroStepper>();
        else if (stepperName == "RK4") _stepper = std::make_shared<gp::RK4Stepper>();
        else throw std::runtime_error("Unknown stepper '" + stepperName + "' (allowed: eulero, RK4)");
        _stepper->setTimeStep(isImaginary ? complex_t(timeStep, 0) : complex_t(0, timeStep));
    }
    void setFunctional(model &m)
    {
        auto func = m.getFunctional();
        _stepper->setFunctional(func);
        _stepper->setDiscretization(func->getDiscretization());
        _stepper->setNComponents(func->getNComponents());
        _stepper->init();
    }
    void setRenormalization(const std::vector<real_t> &normalizations)
    {
        _stepper->setNormalizations(normalizations);
        _stepper->enableReNormalization(true);
    }
    void unsetRenormalization() { _stepper->enableReNormalization(false); }
    void advance(field &f, int nSteps)
    {
        {
            py::gil_scoped_release nogil;
            _stepper->advance(*f.getState(), nSteps, nullptr, nullptr, t);
        }
        tick(nSteps);
    }
    // nSteps, returning per-step norms [nSteps][nC] and energies [nSteps] of the
    // state at the start of each step (fused reductions; addition)
    py::tuple advanceLog(field &f, int nSteps)
    {
        const int nc = f.nComponents();
        py::array_t<double> norms({(py::ssize_t)nSteps, (py::ssize_t)nc});
        py::array_t<double> ens((py::ssize_t)nSteps);
        double *pn = norms.mutable_data(), *pe = ens.mutable_data();
        {
            py::gil_scoped_release nogil;
            _stepper->advance(*f.getState(), nSteps, pn, pe, t);
        }
        tick(nSteps);
        return py::make_tuple(norms, ens);
    }
    real_t getTime() { return t; }

private:
    // t += dt once per step, as the reference's loop does (src/gpCpp.cpp:260-265),
    // so getTime() carries the same rounding as nSteps single additions
    void tick(int nSteps)
    {
        for (int i = 0; i < nSteps; ++i) t += _timeStep;
    }

    std::shared_ptr<gp::stepper> _stepper;
    real_t _timeStep;
    real_t t = 0.0;
};

// src/gpCpp.cpp:289-313: droplet labels of component 0 as an F-ordered int
// array of the field's shape (labelled on the GPU)
class decomposition {
public:
    py::array_t<int> decompose(field &psi, real_t densityCutOff)
    {
        auto state = psi.getState();
        auto d = state->dimensions();
        std::vector<int> labels;
        {
            py::gil_scoped_release nogil;
            gp::decompose::decompose(*state, densityCutOff, labels, 0);
        }
        py::array_t<int> a({(py::ssize_t)d[0], (py::ssize_t)d[1], (py::ssize_t)d[2], (py::ssize_t)d[3]},
                           {(py::ssize_t)sizeof(int), (py::ssize_t)(sizeof(int) * d[0]),
                            (py::ssize_t)(sizeof(int) * d[0] * d[1]), (py::ssize_t)(sizeof(int) * d[0] * d[1] * d[2])});
        std::copy(labels.begin(), labels.end(), a.mutable_data());
        return a;
    }
};

}  // namespace pyInterface

PYBIND11_MODULE(gpCpp, m)
{
    using namespace pyInterface;
    m.doc() = "GP core on one MI355X (HIP kernels via libgp5.so); the reference's gpCpp surface";
    py::class_<geometry>(m, "geometry")
        .def(py::init<std::array<real_t, 3>, std::array<real_t, 3>, std::array<int, 3>>())
        .def("getLeft", &geometry::getLeft)
        .def("getRight", &geometry::getRight)
        .def("getShape", &geometry::getShape);
    py::class_<field>(m, "field")
        .def(py::init<const geometry &, int, std::array<int, 3>, std::array<int, 3>>())
        .def("getLocalShape", &field::getLocalShape)
        .def("getTensor", &field::toPyArray)
        .def("setTensor", &field::setTensor)
        .def("getOffset", &field::getOffset)
        .def("getNorm", &field::getNorm)
        .def("normalize", &field::normalize)
        .def("maxDensity", &field::maxDensity)
        .def("save", &field::save)
        .def("load", &field::load);
    py::class_<model>(m, "model")
        .def("apply", &model::apply)
        .def("energy", &model::energy)
        .def("nComponents", &model::nComponents);
    // the reference binds no constructor; the default one here is an addition
    py::class_<decomposition>(m, "decomposition").def(py::init<>()).def("decompose", &decomposition::decompose);
    py::class_<LHY, model>(m, "LHY")
        .def(py::init<field &>())
        .def(py::init<field &, real_t, real_t>(), py::arg("psi"), py::arg("a"), py::arg("b"));
    py::class_<GP, model>(m, "GP")
        .def(py::init<field &, std::vector<std::vector<real_t>>, std::vector<real_t>,
                      std::vector<std::array<real_t, 3>>>(),
             py::arg("psi"), py::arg("couplings") = std::vector<std::vector<real_t>>{},
             py::arg("masses") = std::vector<real_t>{}, py::arg("omegas") = std::vector<std::array<real_t, 3>>{});
    py::class_<UnlockedLHY, model>(m, "UnlockedLHY").def(py::init<field &, real_t, real_t, real_t>());
    py::class_<timeStepper>(m, "timeStepper")
        .def(py::init<real_t, std::string, bool>())
        .def("setRenormalization", &timeStepper::setRenormalization)
        .def("unsetRenormalization", &timeStepper::unsetRenormalization)
        .def("setFunctional", &timeStepper::setFunctional)
        .def("advance", &timeStepper::advance)
        .def("advanceLog", &timeStepper::advanceLog)
        .def("getTime", &timeStepper::getTime);
}
