// CPU oracle for the orientation-proposal network (gtom_amd/csrc/learn.hip).
//
// TEST INFRASTRUCTURE ONLY: the tests compare the GPU network's logits and
// gradients with this double-precision restatement; the product never links it.
// The network has no counterpart in the reference. With emulate != 0 the
// activations, weights and back-propagated signals are rounded to bf16 at the
// points where the GPU stores them in bf16 (inputs, conv outputs, pooled
// values, the softmax gradient fed to the head's products, the data
// gradients), so GPU and oracle differ only by fp32 accumulation order; with
// emulate == 0 everything is double (the form the finite-difference checks
// pin).
//
// Layouts: tiles [n][32][32][32] (x fastest); activations channels-last
// [S][S][S][16]; parameters as OFF_* below; tap t of a 3^3 kernel is the offset
// (t/9 - 1, (t/3)%3 - 1, t%3 - 1) in (z, y, x).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {
namespace pn {

constexpr int PT = 32, PC = 16, NT = 27, NCLS = 501, NFEAT = 1024;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + PC * NT, OFF_W2 = OFF_B1 + PC, OFF_B2 = OFF_W2 + NT * PC * PC,
              OFF_W3 = OFF_B2 + PC, OFF_B3 = OFF_W3 + NT * PC * PC, OFF_W4 = OFF_B3 + PC,
              OFF_B4 = OFF_W4 + NT * PC * PC, OFF_WF = OFF_B4 + PC, OFF_BF = OFF_WF + NCLS * NFEAT,
              NPARAM = OFF_BF + NCLS;

double bf(double x) {
  float f = (float)x;
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  std::memcpy(&f, &u, 4);
  return f;
}

struct Round {
  bool on;
  double operator()(double x) const { return on ? bf(x) : x; }
};

// weight of input channel ci -> output channel co at tap t
inline double wgt(const double* W, int cin, int t, int co, int ci) {
  return cin == 1 ? W[co * NT + t] : W[(t * PC + co) * PC + ci];
}

bool tap_src(int t, int S, int z, int y, int x, int& v) {
  const int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
  v = (zz * S + yy) * S + xx;
  return zz >= 0 && zz < S && yy >= 0 && yy < S && xx >= 0 && xx < S;
}

// out[v][co] = round(relu?(bias + sum_t sum_ci in[v+off_t][ci] W)), zeroed where mask <= 0
void conv(const std::vector<double>& in, int S, int cin, const double* W, const double* bias, bool relu,
          const std::vector<double>* mask, std::vector<double>& out, Round r) {
  const int V = S * S * S;
  out.assign((size_t)V * PC, 0.0);
  for (int v = 0; v < V; ++v) {
    const int x = v % S, y = (v / S) % S, z = v / (S * S);
    for (int co = 0; co < PC; ++co) {
      double a = bias ? bias[co] : 0.0;
      for (int t = 0; t < NT; ++t) {
        int src;
        if (!tap_src(t, S, z, y, x, src)) continue;
        for (int ci = 0; ci < cin; ++ci) a += in[(size_t)src * cin + ci] * wgt(W, cin, t, co, ci);
      }
      if (relu) a = std::max(a, 0.0);
      if (mask && !((*mask)[(size_t)v * PC + co] > 0)) a = 0.0;
      out[(size_t)v * PC + co] = r(a);
    }
  }
}

// dW[t][co][ci] (or [co][t]) += sum_v dz[v][co] in[v+off_t][ci]; db[co] += sum_v dz[v][co]
void conv_dw(const std::vector<double>& in, int S, int cin, const std::vector<double>& dz, double* dW, double* db) {
  const int V = S * S * S;
  for (int v = 0; v < V; ++v) {
    const int x = v % S, y = (v / S) % S, z = v / (S * S);
    for (int co = 0; co < PC; ++co) {
      const double d = dz[(size_t)v * PC + co];
      if (d == 0.0) continue;
      db[co] += d;
      for (int t = 0; t < NT; ++t) {
        int src;
        if (!tap_src(t, S, z, y, x, src)) continue;
        for (int ci = 0; ci < cin; ++ci) {
          const double g = d * in[(size_t)src * cin + ci];
          if (cin == 1)
            dW[co * NT + t] += g;
          else
            dW[(t * PC + co) * PC + ci] += g;
        }
      }
    }
  }
}

// dX[v][ci] = sum_t sum_co dz[v - off_t][co] W(t, co, ci), rounded, zeroed where mask <= 0.
// g_plant_unflipped (tests only) takes dz[v + off_t] instead, a mistake the
// finite-difference check must catch.
bool g_plant_unflipped = false;
void conv_dx(const std::vector<double>& dz, int S, const double* W, const std::vector<double>* mask,
             std::vector<double>& dx, Round r) {
  const int V = S * S * S;
  dx.assign((size_t)V * PC, 0.0);
  for (int v = 0; v < V; ++v) {
    const int x = v % S, y = (v / S) % S, z = v / (S * S);
    for (int ci = 0; ci < PC; ++ci) {
      double a = 0.0;
      for (int t = 0; t < NT; ++t) {
        int src;
        if (!tap_src(g_plant_unflipped ? t : NT - 1 - t, S, z, y, x, src)) continue;  // v - off_t
        for (int co = 0; co < PC; ++co) a += dz[(size_t)src * PC + co] * W[(t * PC + co) * PC + ci];
      }
      if (mask && !((*mask)[(size_t)v * PC + ci] > 0)) a = 0.0;
      dx[(size_t)v * PC + ci] = r(a);
    }
  }
}

void pool(const std::vector<double>& in, int S, std::vector<double>& out, Round r) {
  const int H = S / 2;
  out.assign((size_t)H * H * H * PC, 0.0);
  for (int z = 0; z < H; ++z)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < H; ++x)
        for (int c = 0; c < PC; ++c) {
          double s = 0;
          for (int d = 0; d < 8; ++d) {
            const int vi = ((2 * z + (d >> 2)) * S + 2 * y + ((d >> 1) & 1)) * S + 2 * x + (d & 1);
            s += in[(size_t)vi * PC + c];
          }
          out[(((size_t)z * H + y) * H + x) * PC + c] = r(s * 0.125);
        }
}

void unpool(const std::vector<double>& dout, int S, const std::vector<double>& act, std::vector<double>& din,
            Round r) {
  const int H = S / 2;
  din.assign((size_t)S * S * S * PC, 0.0);
  for (int z = 0; z < S; ++z)
    for (int y = 0; y < S; ++y)
      for (int x = 0; x < S; ++x)
        for (int c = 0; c < PC; ++c) {
          const size_t i = (((size_t)z * S + y) * S + x) * PC + c;
          const double d = dout[(((size_t)(z / 2) * H + y / 2) * H + x / 2) * PC + c];
          din[i] = act[i] > 0 ? r(d * 0.125) : 0.0;
        }
}

struct Acts {
  std::vector<double> x, a1, p1, a2, p2, a3, a4, feat;
  double logits[NCLS];
};

void forward_one(const double* P, const double* tile, Acts& A, Round r) {
  A.x.resize((size_t)PT * PT * PT);
  for (size_t i = 0; i < A.x.size(); ++i) A.x[i] = r(tile[i]);
  conv(A.x, 32, 1, P + OFF_W1, P + OFF_B1, true, nullptr, A.a1, r);
  pool(A.a1, 32, A.p1, r);
  conv(A.p1, 16, PC, P + OFF_W2, P + OFF_B2, true, nullptr, A.a2, r);
  pool(A.a2, 16, A.p2, r);
  conv(A.p2, 8, PC, P + OFF_W3, P + OFF_B3, true, nullptr, A.a3, r);
  conv(A.a3, 8, PC, P + OFF_W4, P + OFF_B4, true, nullptr, A.a4, r);
  pool(A.a4, 8, A.feat, r);
  for (int c = 0; c < NCLS; ++c) {
    double a = P[OFF_BF + c];
    for (int f = 0; f < NFEAT; ++f) a += A.feat[f] * P[OFF_WF + (size_t)c * NFEAT + f];
    A.logits[c] = a;
  }
}

// parameters as the GPU uses them: weights rounded to bf16 when emulating, biases fp32
std::vector<double> effective(const double* params, Round r) {
  std::vector<double> P(params, params + NPARAM);
  auto rw = [&](int off, int len) {
    for (int i = 0; i < len; ++i) P[off + i] = r(P[off + i]);
  };
  rw(OFF_W1, PC * NT);
  rw(OFF_W2, NT * PC * PC);
  rw(OFF_W3, NT * PC * PC);
  rw(OFF_W4, NT * PC * PC);
  rw(OFF_WF, NCLS * NFEAT);
  if (r.on)
    for (int b : {OFF_B1, OFF_B2, OFF_B3, OFF_B4})
      for (int i = 0; i < PC; ++i) P[b + i] = (float)P[b + i];
  if (r.on)
    for (int i = 0; i < NCLS; ++i) P[OFF_BF + i] = (float)P[OFF_BF + i];
  return P;
}

}  // namespace pn
}  // namespace

extern "C" {

int orc_pnet_param_count() { return pn::NPARAM; }

void orc_pnet_forward_f64(const double* params, const double* tiles, int n, int emulate, double* logits) {
  const pn::Round r{emulate != 0};
  const std::vector<double> P = pn::effective(params, r);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n; ++k) {
    pn::Acts A;
    pn::forward_one(P.data(), tiles + (size_t)k * pn::PT * pn::PT * pn::PT, A, r);
    std::memcpy(logits + (size_t)k * pn::NCLS, A.logits, sizeof(A.logits));
  }
}

// Mean softmax cross-entropy over n tiles and its gradient with respect to every parameter
// (the gradient of the loss at the effective parameters; with emulate the back-propagated
// signals are rounded where the GPU rounds them).
double orc_pnet_gradients_f64(const double* params, const double* tiles, const int* labels, int n, int emulate,
                              double* grads, double* losses) {
  using namespace pn;
  g_plant_unflipped = (emulate & 2) != 0;
  emulate &= 1;
  const Round r{emulate != 0};
  const std::vector<double> P = effective(params, r);
  std::memset(grads, 0, sizeof(double) * NPARAM);
  double total = 0;
  for (int k = 0; k < n; ++k) {
    Acts A;
    forward_one(P.data(), tiles + (size_t)k * PT * PT * PT, A, r);
    double mx = -INFINITY;
    for (int c = 0; c < NCLS; ++c) mx = std::max(mx, A.logits[c]);
    double s = 0;
    for (int c = 0; c < NCLS; ++c) s += std::exp(A.logits[c] - mx);
    const double lse = mx + std::log(s);
    const int lab = labels[k];
    const double loss = lse - A.logits[lab];
    if (losses) losses[k] = loss;
    total += loss / n;
    // dlog (fp32 on the GPU, its bf16 copy feeds the products)
    std::vector<double> dl(NCLS), dlb(NCLS);
    for (int c = 0; c < NCLS; ++c) {
      const double d = (std::exp(A.logits[c] - lse) - (c == lab ? 1.0 : 0.0)) / n;
      dl[c] = r.on ? (double)(float)d : d;
      dlb[c] = r(dl[c]);
      grads[OFF_BF + c] += dl[c];
    }
    std::vector<double> dfeat(NFEAT, 0.0);
    for (int c = 0; c < NCLS; ++c) {
      const double* w = &P[OFF_WF + (size_t)c * NFEAT];
      double* gw = grads + OFF_WF + (size_t)c * NFEAT;
      for (int f = 0; f < NFEAT; ++f) {
        gw[f] += dlb[c] * A.feat[f];
        dfeat[f] += dlb[c] * w[f];
      }
    }
    if (r.on)
      for (double& d : dfeat) d = (double)(float)d;
    std::vector<double> dz4, dz3, dp2, dz2, dp1, dz1;
    unpool(dfeat, 8, A.a4, dz4, r);
    conv_dw(A.a3, 8, PC, dz4, grads + OFF_W4, grads + OFF_B4);
    conv_dx(dz4, 8, &P[OFF_W4], &A.a3, dz3, r);
    conv_dw(A.p2, 8, PC, dz3, grads + OFF_W3, grads + OFF_B3);
    conv_dx(dz3, 8, &P[OFF_W3], nullptr, dp2, r);
    unpool(dp2, 16, A.a2, dz2, r);
    conv_dw(A.p1, 16, PC, dz2, grads + OFF_W2, grads + OFF_B2);
    conv_dx(dz2, 16, &P[OFF_W2], nullptr, dp1, r);
    unpool(dp1, 32, A.a1, dz1, r);
    conv_dw(A.x, 32, 1, dz1, grads + OFF_W1, grads + OFF_B1);
  }
  return total;
}

}  // extern "C"
