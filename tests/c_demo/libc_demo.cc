// A C++ caller of the C interface of the wannierisation library
// (include/w90_library_c.h), written the way callers of wannier90's
// src/wannier90.hh are: options through cset_option_* with Fortran string
// descriptors, cinput_setopt, the k-mesh through cget_nn/cget_nnkp,
// caller-owned M/U/U_opt arrays, then as wannier_prog.F90 orders the steps:
// cdisentangle when num_bands > num_wann (band energies through
// cset_eigval, the projections in U_opt), else cproject_overlap; then
// cwannierise and the result getters.
//
//   libc_demo input.bin output.bin
//
// input.bin (little endian): int32 nw, nb, nk, mp[3], num_iter,
// num_cg_steps, conv_window, dis_num_iter, dis_conv_window; float64
// conv_tol, trial_step, dis_win_min, dis_win_max, dis_froz_min,
// dis_froz_max, dis_mix_ratio, dis_conv_tol (window bounds beyond 1e300
// in size are not set), cell[3][3], kpts[nk][3]; int32 nn; complex128
// M(nb,nb,nn,nk) and A(nb,nw,nk) in Fortran order; when nb > nw float64
// eig(nb,nk).
// output.bin: int32 ierr, nn, nnlist(nk,nn); complex128 U(nw,nw,nk),
// U_opt(nb,nw,nk); float64 centres(3,nw), spreads(nw), omega[4].
#include <complex>
#include <cstdio>
#include <cstring>
#include <vector>

#include "w90_library_c.h"

namespace {

struct FString {
  CFI_CDESC_T(0) desc;
  explicit FString(const char* s) {
    CFI_establish((CFI_cdesc_t*)&desc, (void*)s, CFI_attribute_other, CFI_type_char, std::strlen(s), 0,
                  nullptr);
  }
  CFI_cdesc_t* get() { return (CFI_cdesc_t*)&desc; }
};

template <class T>
bool rd(FILE* f, T* p, size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
    return 2;
  }
  FILE* in = std::fopen(argv[1], "rb");
  if (!in) return 2;
  int nw, nb, nk, mp[3], num_iter, num_cg_steps, conv_window, dis_num_iter, dis_conv_window, nn_in;
  double conv_tol, trial_step, dis[6], cell[3][3];
  bool ok = rd(in, &nw, 1) && rd(in, &nb, 1) && rd(in, &nk, 1) && rd(in, mp, 3) && rd(in, &num_iter, 1) &&
            rd(in, &num_cg_steps, 1) && rd(in, &conv_window, 1) && rd(in, &dis_num_iter, 1) &&
            rd(in, &dis_conv_window, 1) && rd(in, &conv_tol, 1) && rd(in, &trial_step, 1) && rd(in, dis, 6) &&
            rd(in, &cell[0][0], 9);
  const bool disentangle = nb > nw;
  std::vector<double> kpts(3 * (size_t)nk);
  ok = ok && rd(in, kpts.data(), kpts.size()) && rd(in, &nn_in, 1);
  std::vector<std::complex<double>> m((size_t)nb * nb * nn_in * nk), a((size_t)nb * nw * nk);
  std::vector<double> eig(disentangle ? (size_t)nb * nk : 0);
  ok = ok && rd(in, m.data(), m.size()) && rd(in, a.data(), a.size()) && rd(in, eig.data(), eig.size());
  std::fclose(in);
  if (!ok) {
    std::fprintf(stderr, "short input\n");
    return 2;
  }

  void* w90 = w90_create();
  {
    FString k1("num_wann"), k2("num_bands"), k3("num_kpts"), k4("mp_grid"), k5("kpoints"),
        k6("unit_cell_cart"), k7("num_iter"), k8("num_cg_steps"), k9("conv_window"), k10("conv_tol"),
        k11("trial_step");
    cset_option_int(w90, k1.get(), nw);
    cset_option_int(w90, k2.get(), nb);
    cset_option_int(w90, k3.get(), nk);
    int* mpp = mp;
    cset_option_int3(w90, k4.get(), &mpp);
    cset_option_floatxy(w90, k5.get(), kpts.data(), nk, 3);
    double(*cellp)[3] = cell;
    cset_option_float33(w90, k6.get(), &cellp);
    cset_option_int(w90, k7.get(), num_iter);
    cset_option_int(w90, k8.get(), num_cg_steps);
    cset_option_int(w90, k9.get(), conv_window);
    cset_option_float(w90, k10.get(), conv_tol);
    cset_option_float(w90, k11.get(), trial_step);
    if (disentangle) {
      FString d1("dis_num_iter"), d2("dis_conv_window"), d3("dis_mix_ratio"), d4("dis_conv_tol");
      cset_option_int(w90, d1.get(), dis_num_iter);
      cset_option_int(w90, d2.get(), dis_conv_window);
      cset_option_float(w90, d3.get(), dis[4]);
      cset_option_float(w90, d4.get(), dis[5]);
      const char* wkeys[4] = {"dis_win_min", "dis_win_max", "dis_froz_min", "dis_froz_max"};
      for (int i = 0; i < 4; ++i)
        if (dis[i] > -1e300 && dis[i] < 1e300) {
          FString kw(wkeys[i]);
          cset_option_float(w90, kw.get(), dis[i]);
        }
    }
  }
  int ierr = 0;
  FString seed("libc_demo");
  cinput_setopt(w90, seed.get(), &ierr, 0);
  int nn = 0;
  std::vector<int> nnlist;
  if (ierr == 0) {
    cget_nn(w90, &nn);
    nnlist.resize((size_t)nk * nn);
    cget_nnkp(w90, nnlist.data());
    if (nn != nn_in) ierr = 100;
  }
  // without disentanglement U starts as A (replaced by its Lowdin factor)
  // and U_opt is the identity; with it U_opt starts as A and U is output
  std::vector<std::complex<double>> u((size_t)nw * nw * nk, 0.0), uopt((size_t)nb * nw * nk, 0.0);
  if (disentangle) {
    uopt = a;
  } else {
    u = a;
    for (int k = 0; k < nk; ++k)
      for (int i = 0; i < nw; ++i) uopt[(size_t)k * nw * nw + i + (size_t)i * nw] = 1.0;
  }
  std::vector<double> centres(3 * (size_t)nw, 0.0), spreads(nw, 0.0), omega(4, 0.0);
  if (ierr == 0) {
    cset_m_local(w90, m.data());
    cset_u_matrix(w90, u.data());
    cset_u_opt(w90, uopt.data());
    if (disentangle) {
      cset_eigval(w90, eig.data());
      cdisentangle(w90, &ierr);
    } else {
      cproject_overlap(w90, &ierr);
    }
  }
  if (ierr == 0) cwannierise(w90, &ierr);
  if (ierr <= 0) {
    cget_centres(w90, centres.data());
    cget_spreads(w90, spreads.data());
    cget_omega(w90, omega.data());
  }
  w90_delete(w90);

  FILE* out = std::fopen(argv[2], "wb");
  if (!out) return 2;
  std::fwrite(&ierr, sizeof(int), 1, out);
  std::fwrite(&nn, sizeof(int), 1, out);
  std::fwrite(nnlist.data(), sizeof(int), nnlist.size(), out);
  std::fwrite(u.data(), sizeof(std::complex<double>), u.size(), out);
  std::fwrite(uopt.data(), sizeof(std::complex<double>), uopt.size(), out);
  std::fwrite(centres.data(), sizeof(double), centres.size(), out);
  std::fwrite(spreads.data(), sizeof(double), spreads.size(), out);
  std::fwrite(omega.data(), sizeof(double), omega.size(), out);
  std::fclose(out);
  return ierr > 0 ? 1 : 0;
}
