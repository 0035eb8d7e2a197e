// gp5pp.hpp -- C++ facade with the class API of GP5's pseudo-spectral path,
// dispatching to the C ABI of include/gp5.h (HIP kernels on one MI355X).
//
// Class names, setter-style wiring and semantics follow the reference:
//   mesh / domain / discretization       src/geometry.h:9-100
//   fourierTransform(+Creator)           src/fourierTransform.h:14-140
//   operators::laplacian                 src/operators.h:11-26
//   functional, gpFunctional,
//   LHYDropletFunctional,
//   LHYDropletUnlockedFunctional         src/functional.h:9-113
//   externalPotential + kinds            src/externalPotential.h, .cpp:6-82
//   constraint, nullConstraint,
//   phaseConstraint                      src/constraint.h:8-43
//   stepper, euleroStepper, RK4Stepper   src/stepper.h:8-97
//   norm, normalize, initGaussian        src/geometry.h:112-122
// What differs: tensor_t is a device-resident field (host data enters and
// leaves through fromHost/toHost, in the reference's column-major layout) and
// the potential / phase arrays are real FP64 (realTensor_t).  One GPU: the
// processor grid must be {1,1,1}.  Errors throw std::runtime_error with the
// reference's messages where it has one.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <typeinfo>
#include <vector>

#include "../../include/gp5.h"

namespace gp {

using real_t = double;
using complex_t = std::complex<real_t>;
using realDVec_t = std::array<real_t, 3>;
using intDVec_t = std::array<int, 3>;

enum FFT_DIRECTION { FORWARD = 0, BACKWARD = 1 };

inline void check(int rc)
{
    if (rc != GP5_OK) throw std::runtime_error(gp5_last_error());
}

class mesh {
public:
    mesh() : mesh(intDVec_t{0, 0, 0}) {}
    explicit mesh(const intDVec_t &N) : _N(N) {}
    size_t size() const { return (size_t)_N[0] * _N[1] * _N[2]; }
    const intDVec_t &shape() const { return _N; }
    intDVec_t getGlobalOffset() const { return {0, 0, 0}; }
    int getGlobalOffset(int) const { return 0; }
    intDVec_t getNGhosts() const { return {0, 0, 0}; }

private:
    intDVec_t _N;
};

class domain {
public:
    domain() : _left{0, 0, 0}, _right{0, 0, 0} {}
    domain(realDVec_t left, realDVec_t right) : _left(left), _right(right) {}
    realDVec_t getLBox() const { return {_right[0] - _left[0], _right[1] - _left[1], _right[2] - _left[2]}; }
    const realDVec_t &getLeft() const { return _left; }
    const realDVec_t &getRight() const { return _right; }
    void setLeft(const realDVec_t &l) { _left = l; }
    void setRight(const realDVec_t &r) { _right = r; }

private:
    realDVec_t _left, _right;
};

// Owns the device context (grid, FFT plans, stream, scratch) of one
// (shape, box, component count); shared by every tensor of that discretisation.
class device_context {
public:
    device_context(const intDVec_t &shape, const realDVec_t &left, const realDVec_t &right, int nc) : _nc(nc)
    {
        check(gp5_ctx_create(&_ctx, shape.data(), left.data(), right.data(), nc));
    }
    ~device_context() { gp5_ctx_destroy(_ctx); }
    device_context(const device_context &) = delete;
    device_context &operator=(const device_context &) = delete;
    gp5_ctx *get() const { return _ctx; }
    int nComponents() const { return _nc; }

private:
    gp5_ctx *_ctx = nullptr;
    int _nc;
};

inline std::shared_ptr<device_context> shared_context(const intDVec_t &shape, const realDVec_t &left,
                                                     const realDVec_t &right, int nc)
{
    using key_t = std::tuple<intDVec_t, realDVec_t, realDVec_t, int>;
    static std::mutex mtx;
    static std::map<key_t, std::weak_ptr<device_context>> live;
    std::lock_guard<std::mutex> lock(mtx);
    auto &w = live[key_t(shape, left, right, nc)];
    auto p = w.lock();
    if (!p) {
        p = std::make_shared<device_context>(shape, left, right, nc);
        w = p;
    }
    return p;
}

class discretization {
public:
    discretization() = default;
    const std::shared_ptr<mesh> &getLocalMesh() const { return _localMesh; }
    const std::shared_ptr<mesh> &getGlobalMesh() const { return _globalMesh; }
    const std::shared_ptr<domain> &getDomain() const { return _domain; }
    void setLocalMesh(std::shared_ptr<mesh> m) { _localMesh = m; }
    void setGlobalMesh(std::shared_ptr<mesh> m) { _globalMesh = m; }
    void setDomain(std::shared_ptr<domain> d) { _domain = d; }
    // geometry.h:88-90
    real_t getSpaceStep(int d) const { return _domain->getLBox()[d] / _globalMesh->shape()[d]; }
    real_t getFourierStep(int d) const { return 2 * M_PI / _domain->getLBox()[d]; }
    // Every discretisation of the same grid and box shares one device context per
    // component count, so tensors of different fields can meet in one operator call
    // (model.apply(psi1, psi2)) and their work is ordered on one stream.  Calls that
    // share a context must come from one host thread at a time (its scratch is shared).
    std::shared_ptr<device_context> context(int nc)
    {
        auto &p = _ctx[nc];
        if (!p) p = shared_context(_globalMesh->shape(), _domain->getLeft(), _domain->getRight(), nc);
        return p;
    }

private:
    std::shared_ptr<mesh> _globalMesh, _localMesh;
    std::shared_ptr<domain> _domain;
    std::map<int, std::shared_ptr<device_context>> _ctx;
};

// createUniformDiscretization (src/geometry.cpp:56-105), one rank.
inline std::shared_ptr<discretization> createUniformDiscretization(std::shared_ptr<domain> dom,
                                                                   std::shared_ptr<mesh> globalMesh,
                                                                   const intDVec_t &processorGrid = {1, 1, 1})
{
    if (processorGrid[0] != 1 || processorGrid[1] != 1 || processorGrid[2] != 1)
        throw std::runtime_error("one GPU: the processor grid must be {1,1,1}");
    auto d = std::make_shared<discretization>();
    d->setDomain(dom);
    d->setGlobalMesh(globalMesh);
    d->setLocalMesh(std::make_shared<mesh>(globalMesh->shape()));
    return d;
}

// Device buffer of doubles owned by a context.
class device_buffer {
public:
    device_buffer(std::shared_ptr<device_context> ctx, size_t n) : _ctx(ctx), _n(n)
    {
        check(gp5_buf_alloc(ctx->get(), n, &_b));
    }
    ~device_buffer() { gp5_buf_free(_b); }
    device_buffer(const device_buffer &) = delete;
    device_buffer &operator=(const device_buffer &) = delete;
    gp5_buf *get() const { return _b; }
    size_t size() const { return _n; }
    const std::shared_ptr<device_context> &context() const { return _ctx; }

private:
    std::shared_ptr<device_context> _ctx;
    gp5_buf *_b = nullptr;
    size_t _n;
};

// Device-resident complex field (Nx, Ny, Nz, nC), column-major like
// Eigen::Tensor<complex<double>,4> (src/traits.h:12-20).
class tensor_t {
public:
    tensor_t(std::shared_ptr<discretization> discr, int nc) : _discr(discr), _nc(nc)
    {
        const auto &s = discr->getGlobalMesh()->shape();
        _dims = {s[0], s[1], s[2], nc};
        _buf = std::make_shared<device_buffer>(discr->context(nc), 2 * points() * nc);
    }
    std::array<long, 4> dimensions() const { return {_dims[0], _dims[1], _dims[2], _dims[3]}; }
    size_t points() const { return (size_t)_dims[0] * _dims[1] * _dims[2]; }
    size_t size() const { return points() * _nc; }
    int nComponents() const { return _nc; }
    void fromHost(const complex_t *data) { check(gp5_buf_upload(_buf->get(), (const double *)data, 2 * size())); }
    void toHost(complex_t *data) const { check(gp5_buf_download(_buf->get(), (double *)data, 2 * size())); }
    std::vector<complex_t> toHost() const
    {
        std::vector<complex_t> v(size());
        toHost(v.data());
        return v;
    }
    void setZero() { check(gp5_buf_zero(_buf->get())); }
    void copyFrom(const tensor_t &o) { check(gp5_buf_copy(_buf->get(), o._buf->get())); }
    gp5_buf *buf() const { return _buf->get(); }
    gp5_ctx *ctx() const { return _buf->context()->get(); }
    const std::shared_ptr<discretization> &getDiscretization() const { return _discr; }
    // Exchanges the device storage of two fields of one shape (the reference's
    // timeStepper re-points its field at the stepper's buffer, src/gpCpp.cpp:255-269).
    void swap(tensor_t &o)
    {
        if (o._dims != _dims) throw std::runtime_error("swap of fields of different shapes");
        std::swap(_buf, o._buf);
    }

private:
    std::shared_ptr<discretization> _discr;
    int _nc;
    std::array<int, 4> _dims;
    std::shared_ptr<device_buffer> _buf;
};

// out = a - c k and out = a - c (((k1 + 2 k2) + 2 k3) + k4): the Eigen stage
// expressions of src/stepper.cpp:35,50-64 as device passes.
inline void axpy(tensor_t &out, const tensor_t &a, complex_t c, const tensor_t &k)
{
    check(gp5_axpy(out.ctx(), out.buf(), a.buf(), c.real(), c.imag(), k.buf()));
}

inline void rk4Combine(tensor_t &out, const tensor_t &a, complex_t c, const tensor_t &k1, const tensor_t &k2,
                       const tensor_t &k3, const tensor_t &k4)
{
    check(gp5_rk4_combine(out.ctx(), out.buf(), a.buf(), c.real(), c.imag(), k1.buf(), k2.buf(), k3.buf(), k4.buf()));
}

// Device-resident real field (potential, phase): M * nC doubles.
class realTensor_t {
public:
    realTensor_t(std::shared_ptr<discretization> discr, int nc) : _discr(discr), _nc(nc)
    {
        _n = discr->getGlobalMesh()->size() * nc;
        _buf = std::make_shared<device_buffer>(discr->context(nc), _n);
    }
    void fromHost(const real_t *data) { check(gp5_buf_upload(_buf->get(), data, _n)); }
    std::vector<real_t> toHost() const
    {
        std::vector<real_t> v(_n);
        check(gp5_buf_download(_buf->get(), v.data(), _n));
        return v;
    }
    size_t size() const { return _n; }
    int nComponents() const { return _nc; }
    gp5_buf *buf() const { return _buf->get(); }

private:
    std::shared_ptr<discretization> _discr;
    int _nc;
    size_t _n;
    std::shared_ptr<device_buffer> _buf;
};

// Cell-centred coordinates along d (src/geometry.cpp:165-188), host vector.
inline std::vector<real_t> positions(const discretization &discr, int d)
{
    const int n = discr.getGlobalMesh()->shape()[d];
    std::vector<real_t> x(n);
    for (int i = 0; i < n; ++i) x[i] = discr.getDomain()->getLeft()[d] + discr.getSpaceStep(d) * (i + 0.5);
    return x;
}

// Wavenumbers along d in FFT order (src/geometry.cpp:108-152), host vector.
inline std::vector<real_t> momentums(const discretization &discr, int d)
{
    const int n = discr.getGlobalMesh()->shape()[d];
    const real_t dk = discr.getFourierStep(d);
    std::vector<real_t> k(n);
    for (int i = 0; i < n; ++i) {
        if (n % 2 == 0) k[i] = i < n / 2 ? dk * i : (real_t)(i - n) * dk;
        else k[i] = i <= (n - 1) / 2 ? dk * i : (real_t)(i - n) * dk;
    }
    return k;
}

// ---------------------------------------------------------------------------
// transforms
// ---------------------------------------------------------------------------
class fourierTransform {
public:
    virtual ~fourierTransform() = default;
    virtual void apply(tensor_t &source, tensor_t &destination, FFT_DIRECTION dir) = 0;
    virtual std::shared_ptr<discretization> getDiscretizationRealSpace() = 0;
    virtual std::shared_ptr<discretization> getDiscretizationFourierSpace() = 0;
    virtual int getNComponents() const { return _nComponents; }
    virtual void setNComponents(int n) { _nComponents = n; }

private:
    int _nComponents = 1;
};

// Unnormalised 3-D c2c DFT per component on the GPU (the fftwFourierTransform
// counterpart, src/fourierTransform.cpp:129-214): out-of-place with the source
// preserved, k-space in natural order.
class hipFourierTransform : public fourierTransform {
public:
    hipFourierTransform(std::shared_ptr<discretization> discr, int nComponents = 1) : _discr(discr)
    {
        setNComponents(nComponents);
        discr->context(nComponents);  // plans are built with the context
    }
    void apply(tensor_t &source, tensor_t &destination, FFT_DIRECTION dir) override
    {
        if (dir != FORWARD && dir != BACKWARD) throw std::runtime_error("Unkown direction in fourier transform");
        check(gp5_fft(source.ctx(), source.buf(), destination.buf(), dir == FORWARD ? GP5_FORWARD : GP5_BACKWARD));
    }
    std::shared_ptr<discretization> getDiscretizationRealSpace() override { return _discr; }
    std::shared_ptr<discretization> getDiscretizationFourierSpace() override { return _discr; }

private:
    std::shared_ptr<discretization> _discr;
};

// fourierTransformCreator (src/fourierTransform.cpp:228-260): one GPU always
// yields the single-device transform; the ordering is accepted (it only
// permutes the multi-rank P3DFFT layout).
class fourierTransformCreator {
public:
    void setDomain(std::shared_ptr<domain> d) { _domain = d; }
    void setGlobalMesh(std::shared_ptr<mesh> m) { _mesh = m; }
    void setProcessorGrid(const intDVec_t &g) { _grid = g; }
    void setNComponents(int n) { _nc = n; }
    void setOrdering(const intDVec_t &o) { _ordering = o; }
    std::shared_ptr<fourierTransform> create()
    {
        auto discr = createUniformDiscretization(_domain, _mesh, _grid);
        return std::make_shared<hipFourierTransform>(discr, _nc);
    }

private:
    std::shared_ptr<domain> _domain;
    std::shared_ptr<mesh> _mesh;
    intDVec_t _grid{1, 1, 1};
    intDVec_t _ordering{2, 1, 0};
    int _nc = 1;
};

namespace operators {
// src/operators.cpp:7-36: dst = IFFT(-K^2 FFT(src)) / M, one fused pipeline.
class laplacian {
public:
    explicit laplacian(std::shared_ptr<fourierTransform> fft) : _fft(fft) {}
    void apply(tensor_t &source, tensor_t &destination)
    {
        check(gp5_laplacian(source.ctx(), source.buf(), destination.buf()));
    }
    std::shared_ptr<fourierTransform> getFourierTransform() const { return _fft; }

private:
    std::shared_ptr<fourierTransform> _fft;
};
}  // namespace operators

// ---------------------------------------------------------------------------
// functionals
// ---------------------------------------------------------------------------
class functional {
public:
    functional() = default;
    virtual ~functional() = default;
    virtual void apply(tensor_t &fieldOld, tensor_t &fieldNew, real_t time)
    {
        gp5_functional_desc d = desc();
        check(gp5_functional_apply(fieldOld.ctx(), &d, fieldOld.buf(), fieldNew.buf(), time));
    }
    void setLaplacianOperator(std::shared_ptr<operators::laplacian> op) { _lap = op; }
    std::shared_ptr<operators::laplacian> getLaplacianOperator() { return _lap; }
    virtual std::shared_ptr<discretization> getDiscretization() { return _discr; }
    virtual void setDiscretization(std::shared_ptr<discretization> d) { _discr = d; }
    int getNComponents() const { return _nComponents; }
    void setNComponents(int n) { _nComponents = n; }
    void setExternalPotential(std::shared_ptr<realTensor_t> V) { _V = V; }
    std::shared_ptr<realTensor_t> getExternalPotential() const { return _V; }
    virtual void init() {}
    // E[psi] = sum_c int psi_c^* (-1/2 m^-1 lap) psi_c + V rho + e(rho) (extension)
    real_t energy(const tensor_t &psi)
    {
        gp5_functional_desc d = desc();
        double e = 0.0;
        check(gp5_energy(psi.ctx(), &d, psi.buf(), &e));
        return e;
    }
    // The parameters as the C ABI descriptor.  A functional defined outside this
    // header has none: steppers then call its apply() per stage (as the reference's
    // steppers do through the virtual call, src/stepper.cpp:32,47-59).
    virtual gp5_functional_desc desc() const
    {
        throw std::runtime_error("this functional has no device descriptor (it is applied through apply())");
    }

protected:
    gp5_functional_desc base_desc(int kind) const
    {
        gp5_functional_desc d{};
        d.kind = kind;
        d.masses[0] = d.masses[1] = 1.0;
        d.lhy_a = -3.0;
        d.lhy_b = 2.5;
        d.potential = _V ? _V->buf() : nullptr;
        return d;
    }

private:
    std::shared_ptr<operators::laplacian> _lap;
    std::shared_ptr<discretization> _discr;
    int _nComponents = 1;
    std::shared_ptr<realTensor_t> _V;
};

// src/functional.cpp:13-102
class gpFunctional : public functional {
public:
    // couplings[i][j] = g_ij (nComponents x nComponents), as the YAML factory
    // assembles them (src/functional.cpp:154-185)
    void setCouplings(const std::vector<std::vector<real_t>> &g)
    {
        if ((int)g.size() > 2) throw std::runtime_error("couplings support 1 or 2 components");
        for (auto &row : g)
            if (row.size() != g.size()) throw std::runtime_error("Incompatible number of components");
        _g = g;
        _setCouplings = true;
    }
    void setMasses(const std::vector<real_t> &m) { _masses = m; }
    void init() override { _masses.resize(getNComponents(), 1.0); }
    gp5_functional_desc desc() const override
    {
        gp5_functional_desc d = base_desc(GP5_FUNC_GP);
        d.has_couplings = _setCouplings ? 1 : 0;
        for (size_t i = 0; i < _g.size(); ++i)
            for (size_t j = 0; j < _g.size(); ++j) d.couplings[i + 2 * j] = _g[i][j];
        for (size_t c = 0; c < _masses.size() && c < 2; ++c) d.masses[c] = _masses[c];
        return d;
    }

private:
    std::vector<std::vector<real_t>> _g;
    std::vector<real_t> _masses;
    bool _setCouplings = false;
};

// src/functional.cpp:105-122, coefficients a = -3, b = 5/2 unless changed
class LHYDropletFunctional : public functional {
public:
    void setCoefficients(real_t a, real_t b)
    {
        _a = a;
        _b = b;
    }
    gp5_functional_desc desc() const override
    {
        gp5_functional_desc d = base_desc(GP5_FUNC_LHY);
        d.lhy_a = _a;
        d.lhy_b = _b;
        return d;
    }

private:
    real_t _a = -3.0, _b = 2.5;
};

// src/functional.cpp:125-150
class LHYDropletUnlockedFunctional : public functional {
public:
    LHYDropletUnlockedFunctional(real_t alpha, real_t beta, real_t eta) : _alpha(alpha), _beta(beta), _eta(eta) {}
    gp5_functional_desc desc() const override
    {
        gp5_functional_desc d = base_desc(GP5_FUNC_LHY_UNLOCKED);
        d.alpha = _alpha;
        d.beta = _beta;
        d.eta = _eta;
        return d;
    }

private:
    real_t _alpha, _beta, _eta;
};

// ---------------------------------------------------------------------------
// droplet labelling (src/cluster.h:8-18, src/cluster.cpp:17-82) on the GPU:
// labels gets the field's M * nC ints, column-major; returns the droplet count
// ---------------------------------------------------------------------------
namespace decompose {
inline int decompose(const tensor_t &psi, real_t cutOff, std::vector<int> &labels, int component = 0)
{
    labels.assign(psi.size(), 0);
    int n = 0;
    check(gp5_decompose(psi.ctx(), psi.buf(), component, cutOff, labels.data(), &n));
    return n;
}
}  // namespace decompose

// ---------------------------------------------------------------------------
// HDF5 field files (src/io.cpp:56-254): dataset "field", double[2] elements,
// dims (nC, Nz, Ny, Nx); staged through host memory
// ---------------------------------------------------------------------------
inline std::vector<complex_t> loadHost(const std::string &filename, const discretization &discr, int nComponents)
{
    const auto &s = discr.getGlobalMesh()->shape();
    const int shape[3] = {s[0], s[1], s[2]};
    std::vector<complex_t> host((size_t)s[0] * s[1] * s[2] * nComponents);
    check(gp5_h5_read(filename.c_str(), reinterpret_cast<double *>(host.data()), shape, nComponents));
    return host;
}

inline void save(const tensor_t &field, const std::string &filename, const discretization &discr)
{
    const auto &s = discr.getGlobalMesh()->shape();
    const int shape[3] = {s[0], s[1], s[2]};
    auto host = field.toHost();
    check(gp5_h5_write(filename.c_str(), reinterpret_cast<const double *>(host.data()), shape, field.nComponents()));
}

inline tensor_t load(const std::string &filename, std::shared_ptr<discretization> discr, int nComponents)
{
    tensor_t t(discr, nComponents);
    auto host = loadHost(filename, *discr, nComponents);
    t.fromHost(host.data());
    return t;
}

// ---------------------------------------------------------------------------
// external potentials (src/externalPotential.cpp:6-82), evaluated on the host
// once and kept on the device as real FP64
// ---------------------------------------------------------------------------
class externalPotential {
public:
    virtual ~externalPotential() = default;
    // V for component c at grid point (i,j,k)
    virtual void fill(const discretization &discr, int nComponents, std::vector<real_t> &V) const = 0;
    std::shared_ptr<realTensor_t> create(std::shared_ptr<discretization> discr, int nComponents) const
    {
        std::vector<real_t> V(discr->getGlobalMesh()->size() * nComponents, 0.0);
        fill(*discr, nComponents, V);
        auto t = std::make_shared<realTensor_t>(discr, nComponents);
        t->fromHost(V.data());
        return t;
    }
};

class harmonicPotential : public externalPotential {
public:
    explicit harmonicPotential(const std::vector<realDVec_t> &omegas) : _omegas(omegas) {}
    void fill(const discretization &discr, int nc, std::vector<real_t> &V) const override
    {
        if ((int)_omegas.size() < nc) throw std::runtime_error("Incompatible number of components");
        const auto &s = discr.getGlobalMesh()->shape();
        auto x = positions(discr, 0), y = positions(discr, 1), z = positions(discr, 2);
        size_t p = 0;
        for (int c = 0; c < nc; ++c) {
            const auto &w = _omegas[c];
            for (int k = 0; k < s[2]; ++k)
                for (int j = 0; j < s[1]; ++j)
                    for (int i = 0; i < s[0]; ++i, ++p)
                        V[p] = (w[0] * w[0] * x[i] * x[i] + w[1] * w[1] * y[j] * y[j] + w[2] * w[2] * z[k] * z[k]) * 0.5;
        }
    }

private:
    std::vector<realDVec_t> _omegas;
};

class vortexPotential : public externalPotential {
public:
    void fill(const discretization &discr, int nc, std::vector<real_t> &V) const override
    {
        const auto &s = discr.getGlobalMesh()->shape();
        auto x = positions(discr, 0), y = positions(discr, 1);
        size_t p = 0;
        for (int c = 0; c < nc; ++c)
            for (int k = 0; k < s[2]; ++k)
                for (int j = 0; j < s[1]; ++j)
                    for (int i = 0; i < s[0]; ++i, ++p) V[p] = 0.5 / (x[i] * x[i] + y[j] * y[j]);
    }
};

// potentialFromFile (src/externalPotential.cpp:76-82): V read from an HDF5
// field file (gp::load).  The reference keeps V complex; this build adds a real
// V on the device, so a file whose V has a nonzero imaginary part is rejected.
class potentialFromFile : public externalPotential {
public:
    explicit potentialFromFile(std::string filename) : _filename(std::move(filename)) {}
    void fill(const discretization &discr, int nc, std::vector<real_t> &V) const override
    {
        auto z = loadHost(_filename, discr, nc);
        for (size_t p = 0; p < V.size(); ++p) {
            if (z[p].imag() != 0.0)
                throw std::runtime_error(_filename + ": complex potentials are not supported (nonzero imaginary part)");
            V[p] = z[p].real();
        }
    }

private:
    std::string _filename;
};

class sumPotential : public externalPotential {
public:
    explicit sumPotential(std::vector<std::shared_ptr<externalPotential>> pots) : _pots(std::move(pots)) {}
    void fill(const discretization &discr, int nc, std::vector<real_t> &V) const override
    {
        std::vector<real_t> tmp(V.size());
        for (auto &p : _pots) {
            std::fill(tmp.begin(), tmp.end(), 0.0);
            p->fill(discr, nc, tmp);
            for (size_t i = 0; i < V.size(); ++i) V[i] += tmp[i];
        }
    }

private:
    std::vector<std::shared_ptr<externalPotential>> _pots;
};

// ---------------------------------------------------------------------------
// constraints (src/constraint.h:8-43)
// ---------------------------------------------------------------------------
class constraint {
public:
    virtual ~constraint() = default;
    virtual void apply(tensor_t &field) const = 0;
    virtual const realTensor_t *phase() const { return nullptr; }
};

class nullConstraint : public constraint {
public:
    void apply(tensor_t &) const override {}
};

class phaseConstraint : public constraint {
public:
    explicit phaseConstraint(std::shared_ptr<realTensor_t> phase) : _phase(phase) {}
    void apply(tensor_t &field) const override { check(gp5_phase_constraint(field.ctx(), field.buf(), _phase->buf())); }
    const realTensor_t *phase() const override { return _phase.get(); }

private:
    std::shared_ptr<realTensor_t> _phase;
};

// ---------------------------------------------------------------------------
// norm / normalize (src/geometry.cpp:190-237), Gaussian (:241-276)
// ---------------------------------------------------------------------------
inline real_t norm(const tensor_t &field, int c, std::shared_ptr<discretization> = nullptr)
{
    double out = 0.0;
    check(gp5_norm(field.ctx(), field.buf(), c, &out));
    return out;
}

inline void normalize(real_t N, tensor_t &field, int c, std::shared_ptr<discretization> = nullptr)
{
    check(gp5_normalize(field.ctx(), field.buf(), c, N));
}

inline void normalize(const std::vector<int> &N, tensor_t &field, std::shared_ptr<discretization> discr = nullptr)
{
    for (size_t c = 0; c < N.size(); ++c) normalize((real_t)N[c], field, (int)c, discr);
}

inline void initGaussian(const realDVec_t &sigma, std::shared_ptr<discretization> discr, tensor_t &tensor, int comp)
{
    std::vector<complex_t> h = tensor.toHost();
    const auto &s = discr->getGlobalMesh()->shape();
    auto x = positions(*discr, 0), y = positions(*discr, 1), z = positions(*discr, 2);
    const realDVec_t a = {1. / (2 * sigma[0] * sigma[0]), 1. / (2 * sigma[1] * sigma[1]), 1. / (2 * sigma[2] * sigma[2])};
    size_t p = (size_t)comp * s[0] * s[1] * s[2];
    for (int k = 0; k < s[2]; ++k)
        for (int j = 0; j < s[1]; ++j)
            for (int i = 0; i < s[0]; ++i, ++p)
                h[p] = complex_t(std::exp(-(a[0] * x[i] * x[i] + a[1] * y[j] * y[j] + a[2] * z[k] * z[k])), 0);
    tensor.fromHost(h.data());
}

inline void initGaussian(real_t sigma, std::shared_ptr<discretization> discr, tensor_t &tensor, int comp)
{
    initGaussian(realDVec_t{sigma, sigma, sigma}, discr, tensor, comp);
}

// ---------------------------------------------------------------------------
// steppers (src/stepper.h:8-97, src/stepper.cpp:7-81)
//
// With a functional and a constraint of the kinds above, a step runs on the fused
// device path (gp5_advance: every stage is five FFT passes with the functional and
// the stage algebra in the epilogue).  Any other functional or constraint (a class
// derived by the user, as the reference's virtual interfaces allow) is honoured
// through its own apply(): the step then runs stage by stage as the reference
// writes it, with the stage algebra and renormalisation as device passes.
// ---------------------------------------------------------------------------
class stepper {
public:
    stepper() : _constraint(std::make_shared<nullConstraint>()) {}
    virtual ~stepper() = default;
    // One step old -> new (the reference's signature).
    virtual void advance(tensor_t &fieldOld, tensor_t &fieldNew, real_t time)
    {
        if (fused()) {
            fieldNew.copyFrom(fieldOld);
            advance(fieldNew, 1);
        } else {
            stageStep(fieldOld, fieldNew, time);
        }
    }
    // nSteps in place on the device: the hot loop (one call, no host round trip).
    // normLog[s*nC + c] / energyLog[s]: norm and energy at the start of step s
    // (the energy is NaN for a functional without a device descriptor).  time: the
    // time of the first step, handed to a user functional's apply (the fused
    // functionals do not depend on it, as in the reference).
    void advance(tensor_t &field, int nSteps, double *normLog = nullptr, double *energyLog = nullptr,
                 real_t time = 0)
    {
        if (!_func) throw std::runtime_error("stepper has no functional");
        if (fused()) {
            gp5_functional_desc f = _func->desc();
            gp5_stepper_desc s = desc();
            check(gp5_advance(field.ctx(), &f, &s, field.buf(), nSteps, normLog, energyLog));
            return;
        }
        if (!_next || _next->dimensions() != field.dimensions())
            _next = std::make_shared<tensor_t>(field.getDiscretization(), field.nComponents());
        const int nc = field.nComponents();
        real_t t = time;
        for (int st = 0; st < nSteps; ++st) {
            for (int c = 0; c < nc && normLog; ++c) normLog[(size_t)st * nc + c] = norm(field, c);
            if (energyLog) energyLog[st] = std::nan("");
            stageStep(field, *_next, t);
            field.swap(*_next);
            t += std::abs(_timeStep);
        }
    }
    // True when the functional and the constraint are exactly the kinds the
    // fused device path implements.
    bool fused() const
    {
        if (!_func) return false;
        const functional &f = *_func;
        const std::type_info &tf = typeid(f);
        const bool fk = tf == typeid(gpFunctional) || tf == typeid(LHYDropletFunctional) ||
                        tf == typeid(LHYDropletUnlockedFunctional);
        if (!_constraint) return fk;
        const constraint &c = *_constraint;
        const std::type_info &tc = typeid(c);
        return fk && (tc == typeid(nullConstraint) || tc == typeid(phaseConstraint));
    }
    void setTimeStep(complex_t timeStep) { _timeStep = timeStep; }
    complex_t timeStep() const { return _timeStep; }
    void setDiscretization(std::shared_ptr<discretization> d) { _discr = d; }
    std::shared_ptr<discretization> getDiscretization() { return _discr; }
    virtual void init() {}
    virtual void setNormalizations(const std::vector<real_t> &n) { _normalizations = n; }
    virtual const std::vector<real_t> &normalizations() { return _normalizations; }
    std::shared_ptr<functional> getFunctional() { return _func; }
    void setFunctional(std::shared_ptr<functional> f) { _func = f; }
    virtual void setNComponents(int n) { _nComponents = n; }
    int nComponents() { return _nComponents; }
    void enableReNormalization(bool e) { _reNormalize = e; }
    bool reNormalize() const { return _reNormalize; }
    std::shared_ptr<constraint> getConstraint() const { return _constraint; }
    void setConstraint(std::shared_ptr<constraint> c) { _constraint = c; }

protected:
    virtual int kind() const = 0;
    // One step old -> new through the functional's and the constraint's own apply.
    virtual void stageStep(tensor_t &fieldOld, tensor_t &fieldNew, real_t time) = 0;
    // stepper::normalize (src/stepper.cpp:7-22)
    void normalizeField(tensor_t &f)
    {
        if (!_reNormalize) return;
        for (int c = 0; c < f.nComponents(); ++c) {
            if ((size_t)c >= _normalizations.size()) throw std::runtime_error("missing normalization");
            gp::normalize(_normalizations[c], f, c);
        }
    }
    void applyConstraint(tensor_t &f)
    {
        if (_constraint) _constraint->apply(f);
    }
    tensor_t &scratch(std::shared_ptr<tensor_t> &p, const tensor_t &like)
    {
        if (!p || p->dimensions() != like.dimensions())
            p = std::make_shared<tensor_t>(like.getDiscretization(), like.nComponents());
        return *p;
    }
    gp5_stepper_desc desc() const
    {
        if (!_func) throw std::runtime_error("stepper has no functional");
        gp5_stepper_desc s{};
        s.kind = kind();
        s.tau_re = _timeStep.real();
        s.tau_im = _timeStep.imag();
        s.renormalize = _reNormalize ? 1 : 0;
        for (size_t c = 0; c < _normalizations.size() && c < 2; ++c) s.normalizations[c] = _normalizations[c];
        const realTensor_t *ph = _constraint ? _constraint->phase() : nullptr;
        s.phase = ph ? ph->buf() : nullptr;
        return s;
    }

private:
    complex_t _timeStep{0, 0};
    std::shared_ptr<functional> _func;
    int _nComponents = 0;
    std::vector<real_t> _normalizations;
    std::shared_ptr<discretization> _discr;
    bool _reNormalize = false;
    std::shared_ptr<constraint> _constraint;
    std::shared_ptr<tensor_t> _next;  // the other state buffer of the stage-by-stage path
};

class euleroStepper : public stepper {
protected:
    int kind() const override { return GP5_STEP_EULERO; }
    // src/stepper.cpp:30-43
    void stageStep(tensor_t &fieldOld, tensor_t &fieldNew, real_t time) override
    {
        getFunctional()->apply(fieldOld, fieldNew, time);
        axpy(fieldNew, fieldOld, timeStep(), fieldNew);
        applyConstraint(fieldNew);
        normalizeField(fieldNew);
    }
};

class RK4Stepper : public stepper {
protected:
    int kind() const override { return GP5_STEP_RK4; }
    // src/stepper.cpp:46-70 (k1..k4 are allocated on first use, not in init())
    void stageStep(tensor_t &fieldOld, tensor_t &fieldNew, real_t time) override
    {
        tensor_t &k1 = scratch(_k[0], fieldOld), &k2 = scratch(_k[1], fieldOld);
        tensor_t &k3 = scratch(_k[2], fieldOld), &k4 = scratch(_k[3], fieldOld);
        auto f = getFunctional();
        const complex_t tau = timeStep();
        const real_t h = std::abs(tau);
        f->apply(fieldOld, k1, time);
        axpy(fieldNew, fieldOld, tau * 0.5, k1);
        normalizeField(fieldNew);
        f->apply(fieldNew, k2, time + 0.5 * h);
        axpy(fieldNew, fieldOld, tau * 0.5, k2);
        normalizeField(fieldNew);
        f->apply(fieldNew, k3, time + 0.5 * h);
        axpy(fieldNew, fieldOld, tau, k3);
        normalizeField(fieldNew);
        f->apply(fieldNew, k4, time + h);
        rk4Combine(fieldNew, fieldOld, (1. / 6) * tau, k1, k2, k3, k4);
        applyConstraint(fieldNew);
        normalizeField(fieldNew);
    }

private:
    std::shared_ptr<tensor_t> _k[4];
};

}  // namespace gp
