// C++/OpenMP restatement of the reference's OpenMP ("omptarget") kernels for the two benchmark
// chains. TEST INFRASTRUCTURE ONLY: the CPU baseline timed beside the GPU, and a second CPU
// check of the oracle. Nothing in the product links it.
//
// Two forms:
// * "team": the structure the reference's omptarget dialect emits
//   (tensorforge/common/vm/lexic/target_lexic.py:87-168, generators/generator.py:128-196):
//   one `target teams` team per block of `mults` elements, a team-local scratch array for the
//   staged operands, a `parallel` region of lanes*mults threads with bx/ty/tx from the team and
//   thread ids, `omp barrier` between phases. Lane tx computes rows tx, tx+lanes, ... of its
//   element (backend/symbol.py:153-200), accumulating over k in ascending order
//   (backend/instructions/compute/multilinear.py:105-157) and adding the old destination last
//   (multilinear.py:226-247). Only whole blocks are valid (the reference's barrier sits inside
//   the element guard, generator.py:146-152), so numElements must be a multiple of mults.
//   The reference indexes blocks by the team number alone; a host runtime may create fewer
//   teams than num_teams asks for (GCC's host fallback runs one), so here each team walks the
//   blocks bx = team, team + numTeams, ...
// * "flat": one element per loop iteration of a `parallel for`, same per-entry operation order.
//
// Layout: column-major per element, elements contiguous; sparse K in CSR (row_ptr, col, val);
// star matrices packed (column-major pattern order) at the start of each 81-entry slot.
#include <omp.h>
#include <cstddef>
#include <cstring>
#include <vector>

extern "C" {

// C_e += A * B_e  (A b x b shared; B_e, C_e b x q)
void omp_gemm_team(const double* A, const double* B, double* C, int b, int q, long n,
                   int lanes, int mults) {
  const long blocks = n / mults;
  const int threads = lanes * mults;
#pragma omp target teams num_teams(blocks) thread_limit(threads)
  {
    std::vector<double> shr(static_cast<size_t>(mults) * b * q);
    // the host runtime may create fewer teams than asked: each team walks blocks
    for (long bx = omp_get_team_num(); bx < blocks; bx += omp_get_num_teams())
#pragma omp parallel num_threads(threads)
    {
      const int ty = omp_get_thread_num() / lanes;
      const int tx = omp_get_thread_num() % lanes;
      const long e = bx * mults + ty;
      double* lb = &shr[static_cast<size_t>(ty) * b * q];
      const double* gb = B + e * b * q;
      for (int i = tx; i < b * q; i += lanes) lb[i] = gb[i];   // stage B_e
#pragma omp barrier
      const int chunks = (b + lanes - 1) / lanes;
      std::vector<double> acc(static_cast<size_t>(chunks) * q, 0.0);
      for (int k = 0; k < b; ++k)
        for (int c = 0; c < chunks; ++c) {
          const int i = tx + lanes * c;
          if (i >= b) continue;
          const double a = A[i + k * b];
          for (int j = 0; j < q; ++j) acc[c + chunks * j] += a * lb[k + j * b];
        }
      double* gc = C + e * b * q;
      for (int c = 0; c < chunks; ++c) {
        const int i = tx + lanes * c;
        if (i >= b) continue;
        for (int j = 0; j < q; ++j) gc[i + j * b] = acc[c + chunks * j] + gc[i + j * b];
      }
#pragma omp barrier
    }
  }
}

void omp_gemm_flat(const double* A, const double* B, double* C, int b, int q, long n) {
#pragma omp parallel for schedule(static)
  for (long e = 0; e < n; ++e) {
    const double* gb = B + e * b * q;
    double* gc = C + e * b * q;
    for (int j = 0; j < q; ++j)
      for (int i = 0; i < b; ++i) {
        double s = 0.0;
        for (int k = 0; k < b; ++k) s += A[i + k * b] * gb[k + j * b];
        gc[i + j * b] = s + gc[i + j * b];
      }
  }
}

// star product column j of T = Q * S (S packed, pattern given as (row, col) lists in packed order)
static inline void star_product(const double* Qe, const double* Se, const int* srow,
                                const int* scol, int nnz, int b, int q, double* T, int i) {
  for (int j = 0; j < q; ++j) T[i + j * b] = 0.0;
  // k (= star row m) ascending per output entry: packed order is column-major, so within a
  // column the rows ascend
  for (int p = 0; p < nnz; ++p) T[i + scol[p] * b] += Qe[i + srow[p] * b] * Se[p];
}

// Q'_e = sum_d K_d * (Q_e * S_{e,d}); d = 0 overwrites, d > 0 adds the old value last.
void omp_ader_team(const int* const* rowptr, const int* const* col, const double* const* val,
                   const double* Q, const double* const* S, double* Qp, const int* srow,
                   const int* scol, int snnz, int b, int q, long n, int lanes, int mults) {
  const long blocks = n / mults;
  const int threads = lanes * mults;
#pragma omp target teams num_teams(blocks) thread_limit(threads)
  {
    std::vector<double> shr(static_cast<size_t>(mults) * b * q);
    for (long bx = omp_get_team_num(); bx < blocks; bx += omp_get_num_teams())
#pragma omp parallel num_threads(threads)
    {
      const int ty = omp_get_thread_num() / lanes;
      const int tx = omp_get_thread_num() % lanes;
      const long e = bx * mults + ty;
      double* T = &shr[static_cast<size_t>(ty) * b * q];
      const double* Qe = Q + e * b * q;
      double* Qpe = Qp + e * b * q;
      for (int d = 0; d < 3; ++d) {
        const double* Se = S[d] + e * q * q;
        for (int i = tx; i < b; i += lanes) star_product(Qe, Se, srow, scol, snnz, b, q, T, i);
#pragma omp barrier
        for (int i = tx; i < b; i += lanes) {
          for (int j = 0; j < q; ++j) {
            double s = 0.0;
            for (int p = rowptr[d][i]; p < rowptr[d][i + 1]; ++p) s += val[d][p] * T[col[d][p] + j * b];
            Qpe[i + j * b] = d == 0 ? s : s + Qpe[i + j * b];
          }
        }
#pragma omp barrier
      }
    }
  }
}

void omp_ader_flat(const int* const* rowptr, const int* const* col, const double* const* val,
                   const double* Q, const double* const* S, double* Qp, const int* srow,
                   const int* scol, int snnz, int b, int q, long n) {
#pragma omp parallel
  {
    std::vector<double> T(static_cast<size_t>(b) * q);
#pragma omp for schedule(static)
    for (long e = 0; e < n; ++e) {
      const double* Qe = Q + e * b * q;
      double* Qpe = Qp + e * b * q;
      for (int d = 0; d < 3; ++d) {
        const double* Se = S[d] + e * q * q;
        for (int i = 0; i < b; ++i) star_product(Qe, Se, srow, scol, snnz, b, q, T.data(), i);
        for (int j = 0; j < q; ++j)
          for (int i = 0; i < b; ++i) {
            double s = 0.0;
            for (int p = rowptr[d][i]; p < rowptr[d][i + 1]; ++p) s += val[d][p] * T[col[d][p] + j * b];
            Qpe[i + j * b] = d == 0 ? s : s + Qpe[i + j * b];
          }
      }
    }
  }
}

int omp_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
