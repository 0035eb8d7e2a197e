// A functional and a constraint defined by the user, as the reference's virtual
// interfaces allow (src/functional.h:15, src/constraint.h:11): the steppers must
// call their apply() per stage (src/stepper.cpp:30-70) instead of the fused
// device path.  Built and run by tests/test_facade_virtual.py.
//
//   shiftedLHY: the LHY droplet functional plus a constant energy shift v0,
//               H psi = H_LHY psi + v0 psi (what a uniform potential v0 gives);
//   ownPhase:   psi <- |psi| e^{i phi} with the user's own phase array (what
//               phaseConstraint does).
//
// Usage: custom_functional <case> <out.bin>
//   case 0: RK4, imaginary time dt = 1e-3, renormalised to N = 2, ownPhase, 6 steps
//   case 1: Euler, real time dt = 1e-3, 8 steps, no constraint
//   case 2: RK4, real time dt = 1e-3, 5 steps, the built-in functional (fused path)
// The initial field is exp(-r^2/4) (1 + 0.1 i sin x) on 24 x 20 x 16 over [-5,5]x[-5,5]x[-4,4];
// the phase is atan2(y, x + 0.01).  Writes the final field (complex128, column-major).
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gp5pp.hpp"

class shiftedLHY : public gp::functional {
public:
    shiftedLHY(std::shared_ptr<gp::LHYDropletFunctional> inner, double v0) : _inner(inner), _v0(v0) {}
    void apply(gp::tensor_t &fieldOld, gp::tensor_t &fieldNew, gp::real_t time) override
    {
        ++calls;
        _inner->apply(fieldOld, fieldNew, time);
        gp::axpy(fieldNew, fieldNew, gp::complex_t(-_v0, 0.0), fieldOld);  // new += v0 old
    }
    int calls = 0;

private:
    std::shared_ptr<gp::LHYDropletFunctional> _inner;
    double _v0;
};

class ownPhase : public gp::constraint {
public:
    explicit ownPhase(std::shared_ptr<gp::realTensor_t> phi) : _phi(phi) {}
    void apply(gp::tensor_t &field) const override
    {
        ++calls;
        gp::check(gp5_phase_constraint(field.ctx(), field.buf(), _phi->buf()));
    }
    mutable int calls = 0;

private:
    std::shared_ptr<gp::realTensor_t> _phi;
};

int main(int argc, char **argv)
{
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s case out.bin\n", argv[0]);
        return 2;
    }
    const int which = std::atoi(argv[1]);
    try {
        const gp::intDVec_t shape{24, 20, 16};
        gp::fourierTransformCreator fftC;
        fftC.setDomain(std::make_shared<gp::domain>(gp::realDVec_t{-5, -5, -4}, gp::realDVec_t{5, 5, 4}));
        fftC.setGlobalMesh(std::make_shared<gp::mesh>(shape));
        fftC.setNComponents(1);
        auto fft = fftC.create();
        auto discr = fft->getDiscretizationRealSpace();
        auto lap = std::make_shared<gp::operators::laplacian>(fft);

        auto lhy = std::make_shared<gp::LHYDropletFunctional>();
        lhy->setLaplacianOperator(lap);
        lhy->setDiscretization(discr);
        lhy->setNComponents(1);
        auto user = std::make_shared<shiftedLHY>(lhy, 0.35);
        user->setLaplacianOperator(lap);
        user->setDiscretization(discr);
        user->setNComponents(1);

        auto x = gp::positions(*discr, 0), y = gp::positions(*discr, 1), z = gp::positions(*discr, 2);
        std::vector<std::complex<double>> h((size_t)shape[0] * shape[1] * shape[2]);
        std::vector<double> phi(h.size());
        size_t p = 0;
        for (int k = 0; k < shape[2]; ++k)
            for (int j = 0; j < shape[1]; ++j)
                for (int i = 0; i < shape[0]; ++i, ++p) {
                    const double r2 = x[i] * x[i] + y[j] * y[j] + z[k] * z[k];
                    h[p] = std::exp(-r2 / 4.0) * std::complex<double>(1.0, 0.1 * std::sin(x[i]));
                    phi[p] = std::atan2(y[j], x[i] + 0.01);
                }
        gp::tensor_t psi(discr, 1);
        psi.fromHost(h.data());
        auto phase = std::make_shared<gp::realTensor_t>(discr, 1);
        phase->fromHost(phi.data());
        auto con = std::make_shared<ownPhase>(phase);

        std::shared_ptr<gp::stepper> st;
        int nsteps = 0;
        if (which == 0) {
            st = std::make_shared<gp::RK4Stepper>();
            st->setFunctional(user);
            st->setTimeStep({1e-3, 0.0});
            st->setNormalizations({2.0});
            st->enableReNormalization(true);
            st->setConstraint(con);
            nsteps = 6;
        } else if (which == 1) {
            st = std::make_shared<gp::euleroStepper>();
            st->setFunctional(user);
            st->setTimeStep({0.0, 1e-3});
            nsteps = 8;
        } else {
            st = std::make_shared<gp::RK4Stepper>();
            st->setFunctional(lhy);
            st->setTimeStep({0.0, 1e-3});
            nsteps = 5;
        }
        st->setDiscretization(discr);
        st->setNComponents(1);
        st->init();
        st->advance(psi, nsteps);
        psi.toHost(h.data());
        FILE *fp = std::fopen(argv[2], "wb");
        if (!fp) throw std::runtime_error("cannot write the output file");
        std::fwrite(h.data(), sizeof(std::complex<double>), h.size(), fp);
        std::fclose(fp);
        std::printf("fused %d functional_calls %d constraint_calls %d\n", st->fused() ? 1 : 0, user->calls, con->calls);
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 0;
}
