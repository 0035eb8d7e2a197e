// The C++ usage shown for integrators: an LHY droplet advanced with the
// reference-style operator API (gp::RK4Stepper over gp::LHYDropletFunctional
// with a spectral gp::operators::laplacian) on a 64^3 grid.  Built by
// tests/test_capi_load.py (compile and link only when no GPU is present).
#include <complex>
#include <cstdio>
#include <vector>

#include "gp5pp.hpp"

int main()
{
    try {
        gp::fourierTransformCreator fftC;
        fftC.setDomain(std::make_shared<gp::domain>(gp::realDVec_t{-10, -10, -10}, gp::realDVec_t{10, 10, 10}));
        fftC.setGlobalMesh(std::make_shared<gp::mesh>(gp::intDVec_t{64, 64, 64}));
        fftC.setNComponents(1);
        auto fft = fftC.create();
        auto discr = fft->getDiscretizationRealSpace();

        auto func = std::make_shared<gp::LHYDropletFunctional>();
        func->setLaplacianOperator(std::make_shared<gp::operators::laplacian>(fft));
        func->setDiscretization(discr);
        func->setNComponents(1);
        func->init();

        gp::tensor_t psi(discr, 1);
        gp::initGaussian(2.0, discr, psi, 0);
        gp::normalize(500.0, psi, 0);

        gp::RK4Stepper st;
        st.setFunctional(func);
        st.setDiscretization(discr);
        st.setNComponents(1);
        st.setTimeStep({0.0, 1e-3});
        st.init();
        st.advance(psi, 10);

        std::vector<std::complex<double>> host(64 * 64 * 64);
        psi.toHost(host.data());
        std::printf("norm %.12f energy %.12f\n", gp::norm(psi, 0), func->energy(psi));
    } catch (const std::exception &e) {
        std::printf("error: %s\n", e.what());
        return 1;
    }
    return 0;
}
