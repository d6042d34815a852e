// transforms.hpp -- forward/backward spectral transforms of one field on a
// RealFourier / ChebyshevT domain, in the reference's layout and conventions.
#pragma once
#include "runtime.hpp"

namespace tarang {

enum BasisKind : int { kRealFourier = 0, kChebyshevT = 1 };

struct DomainDesc {
    int ndim;
    int kind[3];
    int n[3];
    double lo[3], hi[3];
    double dealias[3];
    int dtype;
};

class TransformPlan {
  public:
    TransformPlan(Context* ctx, const DomainDesc& d);
    const DomainDesc& desc() const { return d_; }
    void coeff_shape(int* shape, int* is_complex) const;
    size_t grid_elems() const;
    size_t coeff_elems() const;  // complex elements
    // device pointers; grid is real (n0, n1, n2), coeff complex (n0/2+1, n1, n2)
    void forward(const void* grid, void* coeff);
    void backward(const void* coeff, void* grid);
    // 3/2-dealiased product of two grid fields (RealFourier-only domains)
    void padded_multiply(const void* ga, const void* gb, void* out);
    // ChebyshevT axes: kChebAuto picks the FFT form where it exists and n >= kChebFFTMinN
    enum { kChebAuto = 0, kChebDense = 1, kChebFFT = 2 };
    void set_cheb_method(int method);
    int cheb_method() const { return cheb_method_; }
    static bool cheb_fft_ok(int n);

  private:
    template <typename T> void forward_t(const void* grid, void* coeff);
    template <typename T> void backward_t(const void* coeff, void* grid);
    template <typename T> void product_t(const void* ga, const void* gb, void* out);
    template <typename T> void axis_fft(void* data, int axis, bool inverse);
    template <typename T> void axis_cheb(const void* in, void* out, int axis, bool forward);  // out may be in
    const void* cheb_matrix(int n, bool forward);
    bool cheb_uses_fft(int n) const;
    // measured on 256 x 256 x n (k_cheb_axis against k_cheb_fft_axis): the FFT form is level at
    // n = 17 (FP64 12.2 against 12.4 us) and faster from n = 33 (16.2 against 20.9 us)
    static constexpr int kChebFFTMinN = 17;
    int cheb_method_ = kChebAuto;

    Context* ctx_;
    DomainDesc d_;
    int cshape_[3];
    DevBuf scratch_[4];
};

}  // namespace tarang
