/*
 * w90_library_c.h — C interface of the Fortran host library (libw90host.so,
 * module w90_library_c), with the names and signatures of wannier90's C
 * interface (src/c_interface.F90, prototypes in src/wannier90.hh), so that
 * callers written against the reference link unchanged.  The Fortran host
 * runs the wannierisation step on an MI355X through libw90gpu.so
 * (include/w90gpu.h).
 *
 * Keyword and seedname strings are Fortran assumed-length character
 * arguments: pass a CFI_cdesc_t built with CFI_establish (type
 * CFI_type_char, elem_len = strlen), as src/wannier90.hh does.  Arrays are
 * the caller's, in Fortran order; they are associated, not copied, and
 * updated in place.  ierr: 0 ok, > 0 error, < 0 warning.
 */
#ifndef W90_LIBRARY_C_H
#define W90_LIBRARY_C_H

#include <ISO_Fortran_binding.h>

#ifdef __cplusplus
extern "C" {
#endif

/* library state object (src/c_interface.F90:10, :17) */
void* w90_create(void);
void w90_delete(void* blob);

/* options (src/c_interface.F90:179-247) */
void cset_option_int(void* blob, CFI_cdesc_t* keyword, int value);
/* int3 and m33 are the address of a pointer to the data (int** / double**),
 * as src/wannier90.hh passes &x */
void cset_option_int3(void* blob, CFI_cdesc_t* keyword, void* int3);
void cset_option_intx(void* blob, CFI_cdesc_t* keyword, int* values, int count);
void cset_option_float(void* blob, CFI_cdesc_t* keyword, double value);
/* 3x3 block, e.g. unit_cell_cart: double[3][3], row i = lattice vector i */
void cset_option_float33(void* blob, CFI_cdesc_t* keyword, void* m33);
/* x rows of y values, e.g. kpoints: double[num_kpts][3], (x, y) = (num_kpts, 3) */
void cset_option_floatxy(void* blob, CFI_cdesc_t* keyword, void* values, int x, int y);

/* apply options / read seedname.win (src/c_interface.F90:49, :70) */
void cinput_setopt(void* blob, CFI_cdesc_t* seedname, int* ierr, int comm);
void cinput_reader(void* blob, int* ierr);

/* caller-owned arrays (src/c_interface.F90:82-127), Fortran order:
 * eigval(num_bands,num_kpts), m(num_bands,num_bands,nntot,num_kpts),
 * u(num_wann,num_wann,num_kpts), u_opt(num_bands,num_wann,num_kpts) */
void cset_eigval(void* blob, double* eigval);
void cset_m_local(void* blob, void* m_matrix);
void cset_u_matrix(void* blob, void* u_matrix);
void cset_u_opt(void* blob, void* u_opt);

/* k-mesh (src/c_interface.F90:129-155): nntot, nnlist(num_kpts,nntot) 1-based */
void cget_nn(void* blob, int* nntot);
void cget_nnkp(void* blob, int* nnlist);

/* steps (src/c_interface.F90:27, :38).  cdisentangle runs dis_main and
 * setup_m_loc (w90_disentangle, src/library_interface.F90:362-420) on the
 * GPU: M (num_bands^2 per (k, b)) through cset_m_local, the projections A
 * in U_opt, the band energies through cset_eigval; it leaves U_opt, U and
 * the num_wann^2 M blocks that cwannierise starts from. */
void cdisentangle(void* blob, int* ierr);
void cwannierise(void* blob, int* ierr);

/* results (src/c_interface.F90:157-177): centres(3,num_wann) Ang,
 * spreads(num_wann) Ang^2 */
void cget_centres(void* blob, void* centres);
void cget_spreads(void* blob, void* spreads);

/* No counterpart in the reference's C interface: the Lowdin projection of
 * w90_project_overlap (src/library_interface.F90:422), with A given in the
 * u_matrix array, for C callers with num_bands = num_wann. */
void cproject_overlap(void* blob, int* ierr);
/* No counterpart in the reference's C interface: the spread decomposition
 * omega[4] = {Omega_I, Omega_D, Omega_OD, Omega} in Ang^2. */
void cget_omega(void* blob, void* omega);

#ifdef __cplusplus
}
#endif

#endif /* W90_LIBRARY_C_H */
