// Dispatcher for oracle/_ref -- test infrastructure only.  arch numbering follows the
// reference's x86 CPU arch list without Eigen (core/src/processor/cpu/CPUProcessor.cpp:17-105):
// 1 = Generic, 4 = FMA, 5 = AVX512.
extern "C" {
int acref_pass_generic(const void*, long, int, int, int, void*, long, int, const float*, const float*);
int acref_pass_fma(const void*, long, int, int, int, void*, long, int, const float*, const float*);
int acref_pass_avx512(const void*, long, int, int, int, void*, long, int, const float*, const float*);

int acref_pass8_generic(const void*, long, int, int, int, void*, long, const float*, const float*, const float*, int);
int acref_pass8_fma(const void*, long, int, int, int, void*, long, const float*, const float*, const float*, int);
int acref_pass8_avx512(const void*, long, int, int, int, void*, long, const float*, const float*, const float*, int);

int acref_artcnn_generic(const void*, long, int, int, int, void*, long, const float*, const float*, int, int);
int acref_artcnn_fma(const void*, long, int, int, int, void*, long, const float*, const float*, int, int);
int acref_artcnn_avx512(const void*, long, int, int, int, void*, long, const float*, const float*, int, int);

int acref_fsrcnnx_generic(const void*, long, int, int, int, void*, long, const float*, const float*, const float*, int, int);
int acref_fsrcnnx_fma(const void*, long, int, int, int, void*, long, const float*, const float*, const float*, int, int);
int acref_fsrcnnx_avx512(const void*, long, int, int, int, void*, long, const float*, const float*, const float*, int, int);

int acref_has_arch(int arch)
{
    __builtin_cpu_init();
    switch (arch) {
    case 1: return 1;
    case 4: return __builtin_cpu_supports("avx") && __builtin_cpu_supports("fma");
    case 5: return __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("fma");
    default: return 0;
    }
}

int acref_pass(int arch, const void* src, long sstride, int w, int h, int stype,
               void* dst, long dstride, int dtype, const float* kernels, const float* biases)
{
    if (!acref_has_arch(arch)) return -95;
    switch (arch) {
    case 1: return acref_pass_generic(src, sstride, w, h, stype, dst, dstride, dtype, kernels, biases);
    case 4: return acref_pass_fma(src, sstride, w, h, stype, dst, dstride, dtype, kernels, biases);
    case 5: return acref_pass_avx512(src, sstride, w, h, stype, dst, dstride, dtype, kernels, biases);
    default: return -22;
    }
}

int acref_pass8(int arch, const void* src, long sstride, int w, int h, int type, void* dst, long dstride,
                const float* kernels, const float* biases, const float* alphas, int blocks)
{
    if (!acref_has_arch(arch)) return -95;
    switch (arch) {
    case 1: return acref_pass8_generic(src, sstride, w, h, type, dst, dstride, kernels, biases, alphas, blocks);
    case 4: return acref_pass8_fma(src, sstride, w, h, type, dst, dstride, kernels, biases, alphas, blocks);
    case 5: return acref_pass8_avx512(src, sstride, w, h, type, dst, dstride, kernels, biases, alphas, blocks);
    default: return -22;
    }
}

int acref_artcnn(int arch, const void* src, long sstride, int w, int h, int type, void* dst, long dstride,
                 const float* kernels, const float* biases, int f, int blocks)
{
    if (!acref_has_arch(arch)) return -95;
    switch (arch) {
    case 1: return acref_artcnn_generic(src, sstride, w, h, type, dst, dstride, kernels, biases, f, blocks);
    case 4: return acref_artcnn_fma(src, sstride, w, h, type, dst, dstride, kernels, biases, f, blocks);
    case 5: return acref_artcnn_avx512(src, sstride, w, h, type, dst, dstride, kernels, biases, f, blocks);
    default: return -22;
    }
}

int acref_fsrcnnx(int arch, const void* src, long sstride, int w, int h, int type, void* dst, long dstride,
                  const float* kernels, const float* biases, const float* alphas, int f, int blocks)
{
    if (!acref_has_arch(arch)) return -95;
    switch (arch) {
    case 1: return acref_fsrcnnx_generic(src, sstride, w, h, type, dst, dstride, kernels, biases, alphas, f, blocks);
    case 4: return acref_fsrcnnx_fma(src, sstride, w, h, type, dst, dstride, kernels, biases, alphas, f, blocks);
    case 5: return acref_fsrcnnx_avx512(src, sstride, w, h, type, dst, dstride, kernels, biases, alphas, f, blocks);
    default: return -22;
    }
}
}
