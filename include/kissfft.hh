/*
 * kissfft.hh -- the reference's C++ class interface (kissfft.hh:16-189) over the GPU
 * library.
 *
 *   kissfft<float> f(nfft, inverse);
 *   f.transform(in, out);                 // complex DFT of nfft points, unnormalised
 *   f.transform(in, out, 0, 1, stride);   // input every stride-th value
 *   f.transform_real(src, dst);           // 2*nfft reals -> dst[0] = (X[0], X[nfft]),
 *                                         // dst[k] = X[k] for 0 < k < nfft
 *   f.assign(nfft2, inverse2);            // as if newly constructed
 *   f.transform_many(in, out, howmany, idist, odist);   // extension: one batched call
 *
 * The transforms run in libkissfft-float (link -lkissfft-float): the class holds that
 * library's cfgs, not a C++ restatement of the algorithm.  Only float is provided (the
 * reference's header also instantiates for double); the recursion parameters `stage` and
 * `fstride` of transform() exist for source compatibility and must keep their defaults.
 * With `inverse`, transform_real() returns the conjugate spectrum (positive exponent), as
 * the reference does.
 */
#ifndef KISSFFT_AMD_CLASS_HH
#define KISSFFT_AMD_CLASS_HH

#include <complex>
#include <cstddef>
#include <cstdlib>
#include <stdexcept>
#include <type_traits>
#include <utility>
#include <vector>

#include "kiss_fft.h"
#include "kiss_fft_batch.h"
#include "kiss_fftr.h"

template <typename scalar_t>
class kissfft
{
    static_assert(std::is_same<scalar_t, float>::value, "kissfft<T>: this library provides float transforms");
    static_assert(sizeof(std::complex<scalar_t>) == sizeof(kiss_fft_cpx), "complex layout");

  public:
    typedef std::complex<scalar_t> cpx_t;

    kissfft(const std::size_t nfft, const bool inverse) : _nfft(nfft), _inverse(inverse) { plan(); }

    kissfft(const kissfft &o) : _nfft(o._nfft), _inverse(o._inverse) { plan(); }

    kissfft(kissfft &&o) noexcept : _nfft(o._nfft), _inverse(o._inverse), _cfg(o._cfg), _rcfg(o._rcfg)
    {
        o._cfg = nullptr;
        o._rcfg = nullptr;
    }

    kissfft &operator=(kissfft o) noexcept
    {
        swap(o);
        return *this;
    }

    ~kissfft()
    {
        std::free(_cfg);
        std::free(_rcfg);
    }

    void swap(kissfft &o) noexcept
    {
        std::swap(_nfft, o._nfft);
        std::swap(_inverse, o._inverse);
        std::swap(_cfg, o._cfg);
        std::swap(_rcfg, o._rcfg);
    }

    /// Changes the length and / or direction (kissfft.hh:55-77).
    void assign(const std::size_t nfft, const bool inverse)
    {
        if (nfft != _nfft || inverse != _inverse) {
            kissfft tmp(nfft, inverse);
            swap(tmp);
        }
    }

    /// Complex DFT of nfft points (kissfft.hh:79-127); output contiguous.
    void transform(const cpx_t *fft_in, cpx_t *fft_out, const std::size_t stage = 0, const std::size_t fstride = 1,
                   const std::size_t in_stride = 1) const
    {
        if (stage != 0 || fstride != 1)
            throw std::invalid_argument("kissfft::transform: stage and fstride are internal (keep 0 and 1)");
        kiss_fft_stride(_cfg, reinterpret_cast<const kiss_fft_cpx *>(fft_in), reinterpret_cast<kiss_fft_cpx *>(fft_out),
                        (int)in_stride);
    }

    /// DFT of 2*nfft real values with the reference's packing (kissfft.hh:129-189).
    void transform_real(const scalar_t *const src, cpx_t *const dst) const
    {
        const std::size_t N = _nfft;
        if (N == 0)
            return;
        std::vector<cpx_t> bins(N + 1);
        kiss_fftr(_rcfg, src, reinterpret_cast<kiss_fft_cpx *>(bins.data()));
        dst[0] = cpx_t(bins[0].real(), bins[N].real());
        for (std::size_t k = 1; k < N; ++k)
            dst[k] = _inverse ? std::conj(bins[k]) : bins[k];
    }

    /// Extension: howmany transforms, idist / odist complex values apart, one GPU call.
    void transform_many(const cpx_t *in, cpx_t *out, const std::size_t howmany, const std::size_t idist,
                        const std::size_t odist) const
    {
        if (kiss_fft_batch(_cfg, reinterpret_cast<const kiss_fft_cpx *>(in), 1, (long long)idist,
                           reinterpret_cast<kiss_fft_cpx *>(out), (long long)odist, (long long)howmany) != 0)
            throw std::runtime_error(kiss_fft_last_error());
    }

    std::size_t nfft() const { return _nfft; }
    bool inverse() const { return _inverse; }

  private:
    void plan()
    {
        _cfg = kiss_fft_alloc((int)_nfft, _inverse ? 1 : 0, nullptr, nullptr);
        _rcfg = kiss_fftr_alloc((int)(2 * _nfft), 0, nullptr, nullptr);
        if (_nfft && (!_cfg || !_rcfg)) {
            std::free(_cfg);
            std::free(_rcfg);
            throw std::invalid_argument("kissfft: cannot plan this length");
        }
    }

    std::size_t _nfft;
    bool _inverse;
    kiss_fft_cfg _cfg = nullptr;
    kiss_fftr_cfg _rcfg = nullptr;
};

#endif
