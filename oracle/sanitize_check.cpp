// Runs the FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer
// (built by `make -C oracle sanitize_check`; test infrastructure only): a
// jittered block mesh with plasticity through reference-mode, relative-
// residual and polar-rotation steps, re-initialisation and the row export.
// Exits 0 when every call returns; the sanitizers abort on the first finding.
#include <cmath>
#include <cstdio>
#include <vector>

extern "C" {
typedef struct {
  double alpha, beta;
  int cg_min_iterations, cg_max_iterations, tol_mode;
  double tol, tiny;
  int rotation_mode;
} orc_params;
typedef struct {
  int iterations, converged;
  double rr0, rr, bb;
} orc_stats;
void *orc_create(int N, int T, const double *coord, const double *model_coord, const double *velocity,
                 const unsigned char *fixed, const int *tets);
void orc_destroy(void *h);
int orc_init(void *h, double young, double poisson, double density, double c_yield, double c_creep,
             double c_max);
int orc_simulate(void *h, double dt, int use_stiffness_warping, const orc_params *P, orc_stats *st);
void orc_default_params(orc_params *P);
void orc_update_original_coord(void *h);
void orc_set_external_forces(void *h, const double *f);
long orc_row_blocks(void *h, int which);
void orc_export_rows(void *h, int which, long *row_ptr, int *cols, double *blocks);
void orc_get_state(void *h, double *x, double *v);
}

int main() {
  const int I = 3, J = 4, K = 2;
  std::vector<double> x;
  std::vector<int> t;
  for (int a = 0; a <= I; ++a)
    for (int b = 0; b <= J; ++b)
      for (int c = 0; c <= K; ++c) {
        const bool in = a > 0 && a < I && b > 0 && b < J && c > 0 && c < K;
        const double j = in ? 0.15 * std::sin(1.7 * a + 2.3 * b + 0.9 * c) : 0.0;
        x.push_back(0.01 * (a + j));
        x.push_back(0.01 * (c - j));
        x.push_back(0.01 * (b + 0.5 * j));
      }
  for (int i = 0; i < I; ++i)
    for (int j = 0; j < J; ++j)
      for (int k = 0; k < K; ++k) {
        int p0 = (i * (J + 1) + j) * (K + 1) + k, p1 = p0 + 1, p3 = ((i + 1) * (J + 1) + j) * (K + 1) + k,
            p2 = p3 + 1, p7 = ((i + 1) * (J + 1) + (j + 1)) * (K + 1) + k, p6 = p7 + 1,
            p4 = (i * (J + 1) + (j + 1)) * (K + 1) + k, p5 = p4 + 1;
        int odd[5][4] = {{p1, p2, p6, p3}, {p3, p6, p4, p7}, {p1, p4, p6, p5}, {p1, p3, p4, p0}, {p1, p6, p4, p3}};
        int even[5][4] = {{p2, p0, p5, p1}, {p2, p7, p0, p3}, {p2, p5, p7, p6}, {p0, p7, p5, p4}, {p2, p0, p7, p5}};
        for (int q = 0; q < 5; ++q)
          for (int c = 0; c < 4; ++c) t.push_back((i + j + k) % 2 ? odd[q][c] : even[q][c]);
      }
  const int N = (int)x.size() / 3, T = (int)t.size() / 4;
  std::vector<unsigned char> fixed(N, 0);
  for (int n = 0; n < N; ++n) fixed[n] = std::fabs(x[3 * n + 2]) < 1e-12;
  void *h = orc_create(N, T, x.data(), x.data(), nullptr, fixed.data(), t.data());
  orc_init(h, 5e5, 0.33, 1000.0, 0.0, 0.5, 0.1);
  std::vector<double> f(3 * N, 0.0);
  for (int n = 0; n < N; ++n) f[3 * n + 1] = fixed[n] ? 0.0 : -2e-3;
  orc_set_external_forces(h, f.data());
  orc_params P;
  orc_stats st;
  orc_default_params(&P);
  for (int s = 0; s < 3; ++s) orc_simulate(h, 0.01, 1, &P, &st);
  P.tol_mode = 1;
  P.tol = 1e-6;
  P.cg_max_iterations = 500;
  P.alpha = 0.1;
  P.beta = 0.001;
  for (int s = 0; s < 3; ++s) orc_simulate(h, 0.01, 1, &P, &st);
  P.rotation_mode = 1;
  orc_simulate(h, 0.01, 1, &P, &st);
  orc_simulate(h, 0.01, 0, &P, &st);
  for (int which = 0; which < 2; ++which) {
    const long nb = orc_row_blocks(h, which);
    std::vector<long> rp(N + 1);
    std::vector<int> cols(nb);
    std::vector<double> blocks(9 * nb);
    orc_export_rows(h, which, rp.data(), cols.data(), blocks.data());
  }
  orc_update_original_coord(h);
  orc_init(h, 5e5, 0.33, 1000.0, 0.0, 0.0, 0.0);
  orc_simulate(h, 0.01, 1, &P, &st);
  std::vector<double> xs(3 * N), vs(3 * N);
  orc_get_state(h, xs.data(), vs.data());
  double sum = 0.0;
  for (double v : xs) sum += v;
  orc_destroy(h);
  std::printf("sanitized oracle run: %d nodes, %d tets, %d iterations, checksum %.6e\n", N, T,
              st.iterations, sum);
  return std::isfinite(sum) ? 0 : 1;
}
