// Test-infrastructure driver for the reference's own CPU translation units.
//
// This file is compiled together with src/evaluation_kernels_cpu.cpp and
// src/pcfg_cpu.cpp of the reference (read in place, never copied) by
// oracle/Makefile.ref into oracle/_ref/libsrref.so.  It only wraps raw
// host pointers into torch::TensorAccessor views (header-only) and calls the
// reference's *_cpu_impl launchers, so that tests can compare the product
// and the oracle against the reference's own CPU algorithm bit for bit.
//
// Nothing in the product links or loads this library.

#include <torch/extension.h>
#include "evaluation_kernels.h"
#include "pcfg_kernels.h"
#include <cstdint>

namespace {

template <typename T, size_t D>
struct View {
    int64_t sizes[D];
    int64_t strides[D];
    T *data;
    View(T *p, std::initializer_list<int64_t> sz) : data(p) {
        size_t i = 0;
        for (auto s : sz) sizes[i++] = s;
        int64_t st = 1;
        for (int j = (int)D - 1; j >= 0; --j) {
            strides[j] = st;
            st *= sizes[j];
        }
    }
    torch::TensorAccessor<T, D> acc() { return torch::TensorAccessor<T, D>(data, sizes, strides); }
};

template <typename T>
void fwd(const T *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const T *C,
         int64_t M, int64_t B, T *cache) {
    View<T, 3> vc(cache, {M, N, B});
    View<int64_t, 2> vo(const_cast<int64_t *>(ops), {M, B});
    View<int64_t, 3> vch(const_cast<int64_t *>(ch), {M, B, 2});
    View<T, 2> vx(const_cast<T *>(X), {N, n_x});
    View<T, 2> vcc(const_cast<T *>(C), {M, B});
    for (int64_t k = 0; k < M; ++k)
        evaluation_forward_step_k_cpu_impl<T>(vc.acc(), vo.acc(), vch.acc(), vx.acc(), vcc.acc(), n_x, k);
}

template <typename T>
int32_t bwd(T *grad_cache, const T *cache, const int64_t *ops, const int64_t *ch, int64_t M, int64_t B,
            int64_t N, int64_t n_x, T *grad_C, T *grad_X) {
    int32_t flag = 0;
    View<T, 3> vg(grad_cache, {M, N, B});
    View<T, 3> vc(const_cast<T *>(cache), {M, N, B});
    View<int64_t, 2> vo(const_cast<int64_t *>(ops), {M, B});
    View<int64_t, 3> vch(const_cast<int64_t *>(ch), {M, B, 2});
    View<T, 2> vgc(grad_C, {M, B});
    View<T, 2> vgx(grad_X, {N, n_x});
    for (int64_t k = M - 1; k >= 0; --k)
        evaluation_backward_step_k_cpu_impl<T>(vg.acc(), vgc.acc(), vgx.acc(), vc.acc(), vo.acc(), vch.acc(), n_x, k,
                                               &flag);
    return flag;
}

template <typename T>
void mfwd(const T *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch, const T *C, int64_t M,
          int64_t B, int64_t K, T *cache) {
    View<T, 4> vc(cache, {M, N, B, K});
    View<int64_t, 2> vo(const_cast<int64_t *>(ops), {M, B});
    View<int64_t, 3> vch(const_cast<int64_t *>(ch), {M, B, 2});
    View<T, 2> vx(const_cast<T *>(X), {N, n_x});
    View<T, 3> vcc(const_cast<T *>(C), {M, B, K});
    for (int64_t k = 0; k < M; ++k)
        evaluation_multiple_forward_step_k_cpu_impl<T>(vc.acc(), vo.acc(), vch.acc(), vx.acc(), vcc.acc(), n_x, k, K);
}

template <typename T>
int32_t mbwd(T *grad_cache, const T *cache, const int64_t *ops, const int64_t *ch, int64_t M, int64_t B,
             int64_t N, int64_t n_x, int64_t K, T *grad_C, T *grad_X) {
    int32_t flag = 0;
    View<T, 4> vg(grad_cache, {M, N, B, K});
    View<T, 4> vc(const_cast<T *>(cache), {M, N, B, K});
    View<int64_t, 2> vo(const_cast<int64_t *>(ops), {M, B});
    View<int64_t, 3> vch(const_cast<int64_t *>(ch), {M, B, 2});
    View<T, 3> vgc(grad_C, {M, B, K});
    View<T, 2> vgx(grad_X, {N, n_x});
    for (int64_t k = M - 1; k >= 0; --k)
        evaluation_multiple_backward_step_k_cpu_impl<T>(vg.acc(), vgc.acc(), vgx.acc(), vc.acc(), vo.acc(), vch.acc(),
                                                        n_x, k, K, &flag);
    return flag;
}

}  // namespace

extern "C" {

int32_t ref_validate(const int64_t *ops, const int64_t *ch, int64_t M, int64_t B) {
    int32_t flag = 0;
    View<int64_t, 2> vo(const_cast<int64_t *>(ops), {M, B});
    View<int64_t, 3> vch(const_cast<int64_t *>(ch), {M, B, 2});
    validate_inputs_cpu_impl(vo.acc(), vch.acc(), &flag);
    return flag;
}

void ref_forward_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                     const float *C, int64_t M, int64_t B, float *cache) {
    fwd<float>(X, N, n_x, ops, ch, C, M, B, cache);
}
void ref_forward_f64(const double *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                     const double *C, int64_t M, int64_t B, double *cache) {
    fwd<double>(X, N, n_x, ops, ch, C, M, B, cache);
}
int32_t ref_backward_f32(float *grad_cache, const float *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                         int64_t B, int64_t N, int64_t n_x, float *grad_C, float *grad_X) {
    return bwd<float>(grad_cache, cache, ops, ch, M, B, N, n_x, grad_C, grad_X);
}
int32_t ref_backward_f64(double *grad_cache, const double *cache, const int64_t *ops, const int64_t *ch, int64_t M,
                         int64_t B, int64_t N, int64_t n_x, double *grad_C, double *grad_X) {
    return bwd<double>(grad_cache, cache, ops, ch, M, B, N, n_x, grad_C, grad_X);
}
void ref_multi_forward_f32(const float *X, int64_t N, int64_t n_x, const int64_t *ops, const int64_t *ch,
                           const float *C, int64_t M, int64_t B, int64_t K, float *cache) {
    mfwd<float>(X, N, n_x, ops, ch, C, M, B, K, cache);
}
int32_t ref_multi_backward_f32(float *grad_cache, const float *cache, const int64_t *ops, const int64_t *ch,
                               int64_t M, int64_t B, int64_t N, int64_t n_x, int64_t K, float *grad_C,
                               float *grad_X) {
    return mbwd<float>(grad_cache, cache, ops, ch, M, B, N, n_x, K, grad_C, grad_X);
}

void ref_sample_infix(const int64_t *rule_lhs, int64_t n_rules, const int64_t *rhs_ptr, const int64_t *rhs_concat,
                      int64_t n_rhs, const int64_t *nt_rule_ptr, int64_t n_sym, const float *cum_probs,
                      int64_t start_id, int64_t max_length, int64_t max_tries, int64_t *output, int64_t *errors,
                      const int64_t *seeds, int64_t B) {
    View<int64_t, 1> a(const_cast<int64_t *>(rule_lhs), {n_rules});
    View<int64_t, 1> b(const_cast<int64_t *>(rhs_ptr), {n_rules + 1});
    View<int64_t, 1> c(const_cast<int64_t *>(rhs_concat), {n_rhs});
    View<int64_t, 1> d(const_cast<int64_t *>(nt_rule_ptr), {n_sym + 1});
    View<float, 1> e(const_cast<float *>(cum_probs), {n_rules});
    View<int64_t, 2> o(output, {B, max_length});
    View<int64_t, 1> er(errors, {B});
    View<int64_t, 1> s(const_cast<int64_t *>(seeds), {B});
    sample_infix_cpu_impl(a.acc(), b.acc(), c.acc(), d.acc(), e.acc(), start_id, max_length, max_tries, o.acc(),
                          er.acc(), s.acc(), B);
}

void ref_infix_to_postfix(const int64_t *prec, int64_t n_prec, const int64_t *expr, int64_t *out, int64_t *errors,
                          int64_t lp, int64_t rp, int64_t B, int64_t M) {
    View<int64_t, 1> p(const_cast<int64_t *>(prec), {n_prec});
    View<int64_t, 2> e(const_cast<int64_t *>(expr), {B, M});
    View<int64_t, 2> o(out, {B, M});
    View<int64_t, 1> er(errors, {B});
    infix_to_postfix_cpu_impl(p.acc(), e.acc(), o.acc(), er.acc(), lp, rp, B, M);
}

void ref_infix_to_prefix(const int64_t *prec, int64_t n_prec, const int64_t *expr, int64_t *out, int64_t *errors,
                         int64_t lp, int64_t rp, int64_t B, int64_t M) {
    View<int64_t, 1> p(const_cast<int64_t *>(prec), {n_prec});
    View<int64_t, 2> e(const_cast<int64_t *>(expr), {B, M});
    View<int64_t, 2> o(out, {B, M});
    View<int64_t, 1> er(errors, {B});
    infix_to_prefix_cpu_impl(p.acc(), e.acc(), o.acc(), er.acc(), lp, rp, B, M);
}

void ref_postfix_to_infix(const int64_t *expr, int64_t *out, int64_t *errors, int64_t lp, int64_t rp, int64_t B,
                          int64_t M, int64_t M_infix) {
    View<int64_t, 2> e(const_cast<int64_t *>(expr), {B, M});
    View<int64_t, 2> o(out, {B, M_infix});
    View<int64_t, 1> er(errors, {B});
    postfix_to_infix_cpu_impl(e.acc(), o.acc(), er.acc(), lp, rp, B, M, M_infix);
}

void ref_prefix_to_infix(const int64_t *expr, int64_t *out, int64_t *errors, int64_t lp, int64_t rp, int64_t B,
                         int64_t M, int64_t M_infix) {
    View<int64_t, 2> e(const_cast<int64_t *>(expr), {B, M});
    View<int64_t, 2> o(out, {B, M_infix});
    View<int64_t, 1> er(errors, {B});
    prefix_to_infix_cpu_impl(e.acc(), o.acc(), er.acc(), lp, rp, B, M, M_infix);
}

void ref_prefix_to_postfix(const int64_t *expr, int64_t *out, int64_t *errors, int64_t B, int64_t M) {
    View<int64_t, 2> e(const_cast<int64_t *>(expr), {B, M});
    View<int64_t, 2> o(out, {B, M});
    View<int64_t, 1> er(errors, {B});
    prefix_to_postfix_cpu_impl(e.acc(), o.acc(), er.acc(), B, M);
}

void ref_postfix_to_prefix(const int64_t *expr, int64_t *out, int64_t *errors, int64_t B, int64_t M) {
    View<int64_t, 2> e(const_cast<int64_t *>(expr), {B, M});
    View<int64_t, 2> o(out, {B, M});
    View<int64_t, 1> er(errors, {B});
    postfix_to_prefix_cpu_impl(e.acc(), o.acc(), er.acc(), B, M);
}

void ref_get_prefix_parent(const int64_t *expr, int64_t *out, int64_t *errors, int64_t B, int64_t M) {
    View<int64_t, 2> e(const_cast<int64_t *>(expr), {B, M});
    View<int64_t, 2> o(out, {B, M});
    View<int64_t, 1> er(errors, {B});
    get_prefix_parent_cpu_impl(e.acc(), o.acc(), er.acc(), B, M);
}

void ref_get_postfix_children(const int64_t *expr, int64_t *out, int64_t *errors, int64_t B, int64_t M) {
    View<int64_t, 2> e(const_cast<int64_t *>(expr), {B, M});
    View<int64_t, 3> o(out, {B, M, 2});
    View<int64_t, 1> er(errors, {B});
    get_postfix_children_cpu_impl(e.acc(), o.acc(), er.acc(), B, M);
}

}  // extern "C"
