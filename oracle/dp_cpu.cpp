// CPU restatement (C++/OpenMP) of the DP se_a energy / force / virial path -- TEST
// INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// may load it: it is the CPU baseline timed beside the GPU and a second, independently derived
// oracle (a hand-written backward instead of autograd) against which the torch oracle is
// checked.  The product path never loads it.
//
// Function -> reference file:line it follows:
//   pair geometry     deepmd_jax/utils.py:20-23,79-82 (orthorhombic minimum image, components
//                     below 1e-15 replaced in the norm, padding masked)
//   switch sr         deepmd_jax/utils.py:29-32
//   env matrix        deepmd_jax/dpmodel.py:37-47, tensor_3to6 deepmd_jax/utils.py:51-52
//   embedding net     deepmd_jax/utils.py:100-125 (tanh layers, repeat-interleave residual)
//   T, G              deepmd_jax/dpmodel.py:50-51,74-78
//   fitting net       deepmd_jax/utils.py:153-166, energy deepmd_jax/dpmodel.py:79-81,100
//   forces / virial   -dE/dx (deepmd_jax/dpmodel.py:106-108) and -dE/dstrain of the pair
//                     vectors (deepmd_jax/md_utils.py:139-159), here by a hand-written
//                     backward: fitting net, G, T, then per pair the embedding net
//                     (recomputed), the env matrix and the switch.
// Templated on the arithmetic type: double pins the torch oracle, float is the FP32 CPU
// baseline.  Atoms are passed type-sorted; each row of the neighbour list holds, per
// neighbour type j, kslots[j] sorted atom indices or -1.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

extern "C" {
struct dpc_model {
  int32_t ntypes, axis, n_embed, n_fit;
  int32_t embed_w[8];  // embedding layer widths (the first layer reads one input)
  int32_t fit_w[8];    // fitting hidden widths (the first layer reads axis x embed_w[last])
  double rcut, nnbrs, out_norm;
  const double* sr_mean;  // (ntypes) per centre type
  const double* sr_std;
  const double* tbias;    // (W)
  const double* ebias;    // (ntypes)
  // per net k = centre type * ntypes + neighbour type, per layer: kernel (in x out) then bias (out)
  const double* embed;
  // per type, per hidden layer: kernel, bias, then dt (out) when the layer's width equals the
  // previous one (l > 0); then the output kernel (last x 1) and bias (1)
  const double* fit;
};
}

namespace {

template <class T>
struct Model {
  int nt, A, ne, nf, W, D;
  std::vector<int> ew, fw;
  T rc, nn, on;
  std::vector<T> mean, sd, tbias, ebias;
  // embedding: per net, per layer: offsets of W, b
  std::vector<std::vector<std::pair<size_t, size_t>>> eoff;
  std::vector<T> ep;
  // fitting: per type, per hidden layer offsets W, b, dt (or npos); output W, b
  struct FitOff {
    std::vector<size_t> w, b, dt;
    size_t ow, ob;
  };
  std::vector<FitOff> foff;
  std::vector<T> fp;

  explicit Model(const dpc_model& m) {
    nt = m.ntypes;
    A = m.axis;
    ne = m.n_embed;
    nf = m.n_fit;
    ew.assign(m.embed_w, m.embed_w + ne);
    fw.assign(m.fit_w, m.fit_w + nf);
    W = ew[ne - 1];
    D = A * W;
    rc = (T)m.rcut;
    nn = (T)m.nnbrs;
    on = (T)m.out_norm;
    for (int t = 0; t < nt; ++t) {
      mean.push_back((T)m.sr_mean[t]);
      sd.push_back((T)m.sr_std[t]);
      ebias.push_back((T)m.ebias[t]);
    }
    for (int w = 0; w < W; ++w) tbias.push_back((T)m.tbias[w]);
    size_t o = 0;
    eoff.resize((size_t)nt * nt);
    for (int k = 0; k < nt * nt; ++k) {
      int in = 1;
      for (int l = 0; l < ne; ++l) {
        const size_t wo = o;
        o += (size_t)in * ew[l];
        const size_t bo = o;
        o += ew[l];
        eoff[k].push_back({wo, bo});
        in = ew[l];
      }
    }
    ep.resize(o);
    for (size_t i = 0; i < o; ++i) ep[i] = (T)m.embed[i];
    o = 0;
    foff.resize(nt);
    for (int t = 0; t < nt; ++t) {
      int in = D;
      for (int l = 0; l < nf; ++l) {
        foff[t].w.push_back(o);
        o += (size_t)in * fw[l];
        foff[t].b.push_back(o);
        o += fw[l];
        if (l > 0 && fw[l] == fw[l - 1]) {
          foff[t].dt.push_back(o);
          o += fw[l];
        } else {
          foff[t].dt.push_back((size_t)-1);
        }
        in = fw[l];
      }
      foff[t].ow = o;
      o += in;
      foff[t].ob = o;
      o += 1;
    }
    fp.resize(o);
    for (size_t i = 0; i < o; ++i) fp[i] = (T)m.fit[i];
  }
};

template <class T>
inline T tnh(T x) {
  return std::tanh(x);
}

// Embedding net forward for one s~: activations of every layer (h[l], width ew[l]); the
// pre-residual tanh values go to th[l] for the backward.
template <class T>
void embed_fwd(const Model<T>& M, int k, T s, std::vector<std::vector<T>>& h, std::vector<std::vector<T>>& th) {
  const T* p = M.ep.data();
  for (int l = 0; l < M.ne; ++l) {
    const int w = M.ew[l];
    const T* Wl = p + M.eoff[k][l].first;
    const T* bl = p + M.eoff[k][l].second;
    if (l == 0) {
      for (int o = 0; o < w; ++o) {
        th[0][o] = tnh(s * Wl[o] + bl[o]);
        h[0][o] = th[0][o];
      }
    } else {
      const int in = M.ew[l - 1], Z = w / in;
      const std::vector<T>& x = h[l - 1];
      for (int o = 0; o < w; ++o) th[l][o] = bl[o];
      for (int i = 0; i < in; ++i) {
        const T xi = x[i];
        const T* row = Wl + (size_t)i * w;
        for (int o = 0; o < w; ++o) th[l][o] += xi * row[o];
      }
      for (int o = 0; o < w; ++o) {
        th[l][o] = tnh(th[l][o]);
        h[l][o] = th[l][o] + x[o / Z];
      }
    }
  }
}

// d(loss)/d s~ given d(loss)/d h[last] (dh, overwritten)
template <class T>
T embed_bwd(const Model<T>& M, int k, const std::vector<std::vector<T>>& th, std::vector<std::vector<T>>& g,
            std::vector<T>& dz) {
  const T* p = M.ep.data();
  for (int l = M.ne - 1; l >= 1; --l) {
    const int w = M.ew[l], in = M.ew[l - 1], Z = w / in;
    const T* Wl = p + M.eoff[k][l].first;
    std::vector<T>& gi = g[l - 1];
    const std::vector<T>& go = g[l];
    for (int o = 0; o < w; ++o) dz[o] = go[o] * (1 - th[l][o] * th[l][o]);
    for (int i = 0; i < in; ++i) gi[i] = 0;
    for (int o = 0; o < w; ++o) gi[o / Z] += go[o];  // residual
    for (int i = 0; i < in; ++i) {
      const T* row = Wl + (size_t)i * w;
      T acc = 0;
      for (int o = 0; o < w; ++o) acc += dz[o] * row[o];
      gi[i] += acc;
    }
  }
  const T* W0 = p + M.eoff[k][0].first;
  T ds = 0;
  for (int o = 0; o < M.ew[0]; ++o) ds += g[0][o] * (1 - th[0][o] * th[0][o]) * W0[o];
  return ds;
}

template <class T>
struct Pair {
  T x[3], r, s, sn, st, xn[3], RX[10];
  int j, k;
  bool valid;
};

template <class T>
inline T minimg(T d, T L) {
  // (d - L/2) mod L - L/2 with a floor-mod
  T y = d - L / 2;
  y = y - L * std::floor(y / L);
  return y - L / 2;
}

template <class T>
int run(const dpc_model* md, int N, const int32_t* types, const double* coord, const double* box3,
        const int32_t* nbr, const int32_t* kslots, double* energy, double* eatom, double* force, double* virial) {
  const Model<T> M(*md);
  int K = 0;
  for (int j = 0; j < M.nt; ++j) K += kslots[j];
  const T L[3] = {(T)box3[0], (T)box3[1], (T)box3[2]};
  const int W = M.W, A = M.A;
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  std::vector<double> fbuf((size_t)nthreads * N * 3, 0.0), vbuf((size_t)nthreads * 9, 0.0);
  std::vector<double> ei(N, 0.0);
  const T sq3 = std::sqrt((T)3), sq2 = std::sqrt((T)2);
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    double* F = fbuf.data() + (size_t)tid * N * 3;
    double* V = vbuf.data() + (size_t)tid * 9;
    std::vector<Pair<T>> pairs(K);
    std::vector<T> Tm(10 * W), dT(10 * W), G(M.D), dG(M.D);
    std::vector<std::vector<T>> h(M.ne), th(M.ne), g(M.ne);
    for (int l = 0; l < M.ne; ++l) {
      h[l].resize(M.ew[l]);
      th[l].resize(M.ew[l]);
      g[l].resize(M.ew[l]);
    }
    int maxf = 1;
    for (int l = 0; l < M.nf; ++l) maxf = maxf > M.fw[l] ? maxf : M.fw[l];
    std::vector<std::vector<T>> y(M.nf, std::vector<T>(maxf)), ty(M.nf, std::vector<T>(maxf));
    std::vector<T> gy(maxf), gy2(maxf), dz(maxf > W ? maxf : W), TW(W), G2(6 * A), dG2(6 * A);
#pragma omp for schedule(dynamic, 16)
    for (int i = 0; i < N; ++i) {
      const int ti = types[i];
      // pair geometry, switch, env matrix
      int slot = 0;
      for (int j = 0; j < M.nt; ++j)
        for (int q = 0; q < kslots[j]; ++q, ++slot) {
          Pair<T>& P = pairs[slot];
          const int a = nbr[(size_t)i * K + slot];
          P.j = a;
          P.k = ti * M.nt + j;
          P.valid = a >= 0;
          for (int d = 0; d < 3; ++d)
            P.x[d] = P.valid ? minimg((T)(coord[(size_t)a * 3 + d] - coord[(size_t)i * 3 + d]), L[d]) : (T)0;
          T r2 = 0;
          for (int d = 0; d < 3; ++d) {
            const T c = std::fabs(P.x[d]) > (T)1e-15 ? P.x[d] : (T)1e-15;
            r2 += c * c;
          }
          P.r = P.valid ? std::sqrt(r2) : (T)0;
          const T t = P.r / M.rc;
          P.s = (P.r > (T)1e-14 && t < 1) ? (1 - 3 * t * t + 2 * t * t * t) / (P.r + (T)1e-15) : (T)0;
          P.sn = P.s / M.sd[ti];
          P.st = (P.s - M.mean[ti]) / M.sd[ti];
          for (int d = 0; d < 3; ++d) P.xn[d] = P.x[d] / (P.r + (T)1e-16);
          P.RX[0] = P.sn;
          for (int c = 0; c < 3; ++c) {
            P.RX[1 + c] = sq3 * P.sn * P.xn[c];
            P.RX[4 + c] = 3 * P.sn * (P.xn[c] * P.xn[c] - (T)1 / 3);
            P.RX[7 + c] = 3 * P.sn * sq2 * P.xn[c] * P.xn[(c + 2) % 3];
          }
        }
      // T = sum over pairs RX (x) h / Nnbrs
      std::fill(Tm.begin(), Tm.end(), (T)0);
      for (int q = 0; q < K; ++q) {
        Pair<T>& P = pairs[q];
        if (!P.valid || P.s == 0) continue;  // padding and pairs beyond rcut: RX = 0
        embed_fwd(M, P.k, P.st, h, th);
        const std::vector<T>& hw = h[M.ne - 1];
        for (int rr = 0; rr < 10; ++rr) {
          const T c = P.RX[rr] / M.nn;
          if (c == 0) continue;
          T* Tr = Tm.data() + rr * W;
          for (int w = 0; w < W; ++w) Tr[w] += c * hw[w];
        }
      }
      // G[a][w] = TW[a] TW[w] + sum_c T3[c][a] T3[c][w] + sum_k G2[k][a] T6[k][w]
      for (int w = 0; w < W; ++w) TW[w] = Tm[w] + M.tbias[w];
      const T* T3 = Tm.data() + W;
      const T* T6 = Tm.data() + 4 * W;
      for (int a = 0; a < A; ++a)
        for (int c = 0; c < 3; ++c) {
          const T u = T3[c * W + A + a], v = T3[((c + 2) % 3) * W + A + a];
          G2[c * A + a] = u * u + T6[c * W + A + a];
          G2[(3 + c) * A + a] = sq2 * u * v + T6[(3 + c) * W + A + a];
        }
      for (int a = 0; a < A; ++a)
        for (int w = 0; w < W; ++w) {
          T v = TW[a] * TW[w];
          for (int c = 0; c < 3; ++c) v += T3[c * W + a] * T3[c * W + w];
          for (int kk = 0; kk < 6; ++kk) v += G2[kk * A + a] * T6[kk * W + w];
          G[(size_t)a * W + w] = v;
        }
      // fitting forward
      const typename Model<T>::FitOff& fo = M.foff[ti];
      const T* fp = M.fp.data();
      int in = M.D;
      const T* x = G.data();
      for (int l = 0; l < M.nf; ++l) {
        const int w = M.fw[l];
        const T* Wl = fp + fo.w[l];
        const T* bl = fp + fo.b[l];
        std::vector<T>& z = ty[l];
        for (int o = 0; o < w; ++o) z[o] = bl[o];
        for (int u = 0; u < in; ++u) {
          const T xu = x[u];
          const T* row = Wl + (size_t)u * w;
          for (int o = 0; o < w; ++o) z[o] += xu * row[o];
        }
        for (int o = 0; o < w; ++o) {
          z[o] = tnh(z[o]);
          y[l][o] = fo.dt[l] != (size_t)-1 ? z[o] * fp[fo.dt[l] + o] + x[o] : z[o];
        }
        x = y[l].data();
        in = w;
      }
      T e = fp[fo.ob];
      for (int u = 0; u < in; ++u) e += x[u] * fp[fo.ow + u];
      ei[i] = (double)((e + M.ebias[ti]) * M.on);
      // fitting backward: dE/dG
      for (int u = 0; u < in; ++u) gy[u] = M.on * fp[fo.ow + u];
      for (int l = M.nf - 1; l >= 0; --l) {
        const int w = M.fw[l], win = l == 0 ? M.D : M.fw[l - 1];
        const T* Wl = fp + fo.w[l];
        const bool res = fo.dt[l] != (size_t)-1;
        // dz = dy * dt * (1 - tanh^2) (residual layers) or dy * (1 - tanh^2)
        for (int o = 0; o < w; ++o) {
          const T tz = ty[l][o];
          dz[o] = gy[o] * (res ? fp[fo.dt[l] + o] : (T)1) * (1 - tz * tz);
        }
        T* out = l == 0 ? dG.data() : gy2.data();
        for (int u = 0; u < win; ++u) {
          const T* row = Wl + (size_t)u * w;
          T acc = res ? gy[u] : (T)0;
          for (int o = 0; o < w; ++o) acc += dz[o] * row[o];
          out[u] = acc;
        }
        if (l > 0)
          for (int u = 0; u < win; ++u) gy[u] = gy2[u];
      }
      // dE/dT from G
      std::fill(dT.begin(), dT.end(), (T)0);
      T* dTW = dT.data();
      T* dT3 = dT.data() + W;
      T* dT6 = dT.data() + 4 * W;
      std::fill(dG2.begin(), dG2.end(), (T)0);
      for (int a = 0; a < A; ++a)
        for (int w = 0; w < W; ++w) {
          const T gg = dG[(size_t)a * W + w];
          dTW[w] += gg * TW[a];
          dTW[a] += gg * TW[w];
          for (int c = 0; c < 3; ++c) {
            dT3[c * W + w] += gg * T3[c * W + a];
            dT3[c * W + a] += gg * T3[c * W + w];
          }
          for (int kk = 0; kk < 6; ++kk) {
            dT6[kk * W + w] += gg * G2[kk * A + a];
            dG2[kk * A + a] += gg * T6[kk * W + w];
          }
        }
      for (int a = 0; a < A; ++a)
        for (int c = 0; c < 3; ++c) {
          const int cp = (c + 2) % 3;
          const T u = T3[c * W + A + a], v = T3[cp * W + A + a];
          dT6[c * W + A + a] += dG2[c * A + a];
          dT6[(3 + c) * W + A + a] += dG2[(3 + c) * A + a];
          dT3[c * W + A + a] += dG2[c * A + a] * 2 * u + dG2[(3 + c) * A + a] * sq2 * v;
          dT3[cp * W + A + a] += dG2[(3 + c) * A + a] * sq2 * u;
        }
      // per pair: dE/dRX, dE/dh -> dE/ds~ (embedding backward), then dE/dx
      for (int q = 0; q < K; ++q) {
        Pair<T>& P = pairs[q];
        if (!P.valid || P.s == 0) continue;
        embed_fwd(M, P.k, P.st, h, th);
        const std::vector<T>& hw = h[M.ne - 1];
        T dRX[10];
        std::vector<T>& gh = g[M.ne - 1];
        for (int w = 0; w < W; ++w) gh[w] = 0;
        for (int rr = 0; rr < 10; ++rr) {
          const T* dTr = dT.data() + rr * W;
          T acc = 0;
          const T c = P.RX[rr] / M.nn;
          for (int w = 0; w < W; ++w) {
            acc += dTr[w] * hw[w];
            gh[w] += c * dTr[w];
          }
          dRX[rr] = acc / M.nn;
        }
        const T dst = embed_bwd(M, P.k, th, g, dz);
        // RX in terms of sn and xn
        T dsn = dRX[0];
        T dxn[3] = {0, 0, 0};
        for (int c = 0; c < 3; ++c) {
          const int cp = (c + 2) % 3;
          dsn += dRX[1 + c] * sq3 * P.xn[c] + dRX[4 + c] * 3 * (P.xn[c] * P.xn[c] - (T)1 / 3) +
                 dRX[7 + c] * 3 * sq2 * P.xn[c] * P.xn[cp];
          dxn[c] += dRX[1 + c] * sq3 * P.sn + dRX[4 + c] * 6 * P.sn * P.xn[c] + dRX[7 + c] * 3 * sq2 * P.sn * P.xn[cp];
          dxn[cp] += dRX[7 + c] * 3 * sq2 * P.sn * P.xn[c];
        }
        const T ds = dsn / M.sd[ti] + dst / M.sd[ti];
        const T r = P.r, t = r / M.rc, re = r + (T)1e-15;
        const T dsdr = ((-6 * t + 6 * t * t) / M.rc * re - (1 - 3 * t * t + 2 * t * t * t)) / (re * re);
        const T ri = r + (T)1e-16;
        T xdot = 0;
        for (int c = 0; c < 3; ++c) xdot += dxn[c] * P.x[c];
        T gx[3];
        for (int d = 0; d < 3; ++d)
          gx[d] = ds * dsdr * P.x[d] / r + dxn[d] / ri - xdot * P.x[d] / (r * ri * ri);
        for (int d = 0; d < 3; ++d) {
          F[(size_t)P.j * 3 + d] -= (double)gx[d];
          F[(size_t)i * 3 + d] += (double)gx[d];
        }
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) V[a * 3 + b] -= (double)(P.x[a] * gx[b]);
      }
    }
  }
  double E = 0;
  for (int i = 0; i < N; ++i) E += ei[i];
  *energy = E;
  if (eatom)
    for (int i = 0; i < N; ++i) eatom[i] = ei[i];
  for (size_t k = 0; k < (size_t)N * 3; ++k) {
    double s = 0;
    for (int t = 0; t < nthreads; ++t) s += fbuf[(size_t)t * N * 3 + k];
    force[k] = s;
  }
  if (virial)
    for (int c = 0; c < 9; ++c) {
      double s = 0;
      for (int t = 0; t < nthreads; ++t) s += vbuf[(size_t)t * 9 + c];
      virial[c] = s;
    }
  return 0;
}

}  // namespace

extern "C" {
// E, per-atom energies (may be NULL), forces (N x 3) and virial (3 x 3, may be NULL) of one
// orthorhombic frame; atoms type-sorted, nbr (N x sum kslots) sorted indices or -1.
int dpc_energy_force_f64(const dpc_model* m, int32_t N, const int32_t* types, const double* coord,
                         const double* box3, const int32_t* nbr, const int32_t* kslots, double* energy,
                         double* eatom, double* force, double* virial) {
  return run<double>(m, N, types, coord, box3, nbr, kslots, energy, eatom, force, virial);
}
int dpc_energy_force_f32(const dpc_model* m, int32_t N, const int32_t* types, const double* coord,
                         const double* box3, const int32_t* nbr, const int32_t* kslots, double* energy,
                         double* eatom, double* force, double* virial) {
  return run<float>(m, N, types, coord, box3, nbr, kslots, energy, eatom, force, virial);
}
int dpc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
}
