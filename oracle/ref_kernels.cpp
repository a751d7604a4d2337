// Driver for the reference's OWN kernel headers — test infrastructure only.
//
// Compiled by oracle/Makefile against the reference sources where they lie
// (-I <reference>/src), once with -DREF_INTRIN (kernels/intrin, the default
// AVX2 build of the reference, cmake/options.cmake:27,93) and once with
// -DREF_NOINTRIN (kernels/nointrin), into oracle/_ref/libref_{intrin,nointrin}.so.
// This file contains no reference code: it owns a state vector of the
// reference's own StateVector/Matrix types and dispatches exactly as
// SimulatorMPI::Run does (SimulatorMPI.cpp:470-514): positions passed high to
// low, an enclosing `#pragma omp parallel`, diag vs dense selection.
#include <complex>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "simulator-mpi/alignedallocator.hpp"

#if defined(REF_INTRIN)
#     include "simulator-mpi/kernels/intrin/kernels.hpp"
#     include "simulator-mpi/kernels/intrin/kernels_diag.hpp"
using intrin::kernel_core;
using intrin::kernel_core_diag;
using intrin::kernelK;
using intrin::kernelK_diag1;
#elif defined(REF_NOINTRIN)
#     include "simulator-mpi/kernels/nointrin/kernels.hpp"
#     include "simulator-mpi/kernels/nointrin/kernels_diag.hpp"
using nointrin::kernel_core;
using nointrin::kernel_core_diag;
using nointrin::kernelK;
using nointrin::kernelK_diag1;
#else
#     error "define REF_INTRIN or REF_NOINTRIN"
#endif

#ifdef _OPENMP
#     include <omp.h>
#endif

namespace {
using Complex = std::complex<double>;

// Allocator of the harness's state vector: 64-byte aligned like the
// reference's StateVector (SimulatorMPI.hpp), but resize() leaves the
// elements untouched, so that ref_state_fill's OpenMP threads touch the
// pages first and each thread's static share of the state lands in its own
// NUMA node, which is how a multi-socket host runs a memory-bound pass at
// full bandwidth.  (The reference zero-fills from one thread.)
template <class T>
struct FirstTouchAlloc {
     using value_type = T;
     FirstTouchAlloc() noexcept {}
     template <class U>
     FirstTouchAlloc(const FirstTouchAlloc<U>&) noexcept
     {}
     T* allocate(size_t n)
     {
          void* p = nullptr;
          if (posix_memalign(&p, 64, n * sizeof(T))) throw std::bad_alloc();
          return static_cast<T*>(p);
     }
     void deallocate(T* p, size_t) noexcept { std::free(p); }
     template <class U>
     void construct(U*) noexcept
     {}
     template <class U, class A0, class... A>
     void construct(U* p, A0&& a0, A&&... a)
     {
          ::new (static_cast<void*>(p)) U(std::forward<A0>(a0), std::forward<A>(a)...);
     }
     template <class U>
     bool operator==(const FirstTouchAlloc<U>&) const noexcept
     {
          return true;
     }
     template <class U>
     bool operator!=(const FirstTouchAlloc<U>&) const noexcept
     {
          return false;
     }
};

using StateVector = std::vector<Complex, FirstTouchAlloc<Complex>>;
using Matrix = std::vector<std::vector<Complex, aligned_allocator<Complex, 64>>>;

struct RefState {
     StateVector vec;
};

Matrix to_matrix(const double* m, size_t dim)
{
     Matrix M(dim, std::vector<Complex, aligned_allocator<Complex, 64>>(dim));
     for (size_t i = 0; i < dim; ++i)
          for (size_t j = 0; j < dim; ++j)
               M[i][j] = Complex(m[2 * (i * dim + j)], m[2 * (i * dim + j) + 1]);
     return M;
}
}  // namespace

extern "C" {

// A state of n_amps amplitudes, zero, first touched by the OpenMP threads.
void* ref_state_new(uint64_t n_amps)
{
     auto* s = new RefState();
     s->vec.resize(n_amps);
     Complex* v = s->vec.data();
     const int64_t n = static_cast<int64_t>(n_amps);
#pragma omp parallel for schedule(static)
     for (int64_t i = 0; i < n; ++i) v[i] = Complex(0.0, 0.0);
     return s;
}

// Every amplitude = (re, im), in parallel with the same static schedule.
void ref_state_fill(void* s, double re, double im)
{
     StateVector& v = static_cast<RefState*>(s)->vec;
     const int64_t n = static_cast<int64_t>(v.size());
#pragma omp parallel for schedule(static)
     for (int64_t i = 0; i < n; ++i) v[i] = Complex(re, im);
}
void ref_state_free(void* s) { delete static_cast<RefState*>(s); }
double* ref_state_data(void* s) { return reinterpret_cast<double*>(static_cast<RefState*>(s)->vec.data()); }

int ref_num_threads()
{
#ifdef _OPENMP
     return omp_get_max_threads();
#else
     return 1;
#endif
}

// Apply one fused matrix exactly as SimulatorMPI::Run does.  pos[l] is the
// bit position of matrix bit l (ids_pos in Run).  k == 0 applies m[0][0] as a
// global factor when it differs from 1.
int ref_apply(void* s, const double* m, int k, const uint32_t* pos, uint64_t ctrl_mask, int diag)
{
     StateVector& vec_ = static_cast<RefState*>(s)->vec;
     Matrix mm = to_matrix(m, size_t(1) << k);
     const uint32_t* p = pos;
     switch (k) {
          case 0:
               if (mm[0][0] != Complex(1.0, 0.0)) kernelK_diag1<StateVector, Complex>(vec_, mm[0][0]);
               break;
          case 1:
#pragma omp parallel
               diag ? kernelK<StateVector, Matrix, kernel_core_diag>(vec_, p[0], mm, ctrl_mask)
                    : kernelK<StateVector, Matrix, kernel_core>(vec_, p[0], mm, ctrl_mask);
               break;
          case 2:
#pragma omp parallel
               diag ? kernelK<StateVector, Matrix, kernel_core_diag>(vec_, p[1], p[0], mm, ctrl_mask)
                    : kernelK<StateVector, Matrix, kernel_core>(vec_, p[1], p[0], mm, ctrl_mask);
               break;
          case 3:
#pragma omp parallel
               diag ? kernelK<StateVector, Matrix, kernel_core_diag>(vec_, p[2], p[1], p[0], mm,
                                                                     ctrl_mask)
                    : kernelK<StateVector, Matrix, kernel_core>(vec_, p[2], p[1], p[0], mm,
                                                                ctrl_mask);
               break;
          case 4:
#pragma omp parallel
               diag ? kernelK<StateVector, Matrix, kernel_core_diag>(vec_, p[3], p[2], p[1], p[0],
                                                                     mm, ctrl_mask)
                    : kernelK<StateVector, Matrix, kernel_core>(vec_, p[3], p[2], p[1], p[0], mm,
                                                                ctrl_mask);
               break;
          case 5:
#pragma omp parallel
               diag ? kernelK<StateVector, Matrix, kernel_core_diag>(vec_, p[4], p[3], p[2], p[1],
                                                                     p[0], mm, ctrl_mask)
                    : kernelK<StateVector, Matrix, kernel_core>(vec_, p[4], p[3], p[2], p[1], p[0],
                                                                mm, ctrl_mask);
               break;
          default:
               return 1;
     }
     return 0;
}

}  // extern "C"
