/* Readers for the reference's HDF5 input files (libafq_h5.so, host only, HDF5 1.10 C API).
 *
 * The reference reads its Hamiltonian and trial files with C++ readers built on its h5 wrapper:
 *   - RealDenseHamiltonian (RealDenseHamiltonian.cpp:76-136) and read_energy_offset
 *     (hdf5_helpers.hpp:84-108) for /Hamiltonian: hcore [nspin][npol][N][npol][N],
 *     DenseFactorized/L [nspin][1][N][1][N][M] (real), E0 = nuclear_energy +
 *     frozen_core_energy - nelec * madelung_constant (each attribute optional);
 *   - read_wavefunction_info (readWfn.cpp:86-110), read_nomsd_wavefunction (readWfn.h:90-156),
 *     read_ph_wavefunction_hdf (readWfn.cpp:148-214) for /Wavefunction/{NOMSD,PHMSD}: spin_type,
 *     ci_coeffs, PsiT_k sparse matrices (HDF2CSR, csr_utils.hpp:353-420: `shape`,
 *     `row_pointers`, `column_indices`, `values`, or the legacy `dims`, `pointers_begin_`,
 *     `pointers_end_`, `jdata_`, `data_`), occa / occb and number_of_orbitals.
 * These entry points replace those readers for the device path: they hand plain row-major
 * arrays to afq_ham_create / afq_phmsd_create.  Complex values are interleaved (re, im) doubles.
 *
 * Every function returns AFQ_OK (0) or a negative afq_status (afq.h); afq_h5_last_error()
 * describes the failure (missing file, group or dataset, or a shape that disagrees).
 */
#ifndef AFQ_H5_H
#define AFQ_H5_H

#ifdef __cplusplus
extern "C" {
#endif

enum { AFQ_H5_NOMSD = 0, AFQ_H5_PHMSD = 1 };
enum { AFQ_H5_CLOSED = 0, AFQ_H5_COLLINEAR = 1, AFQ_H5_NONCOLLINEAR = 2 };

/* Text of the last error of this thread. */
const char* afq_h5_last_error(void);

/* /Hamiltonian extents: nspin of hcore, orbitals N, Cholesky vectors M, nspin of L. */
int afq_h5_hamiltonian_shape(const char* path, int* nspin, int* N, int* M, int* nspin_L);

/* hcore [nspin][N][N] and L as [nspin_L][N][N][M] (the file's [nspin_L][1][N][1][N][M], nspin_L
 * 1 or nspin); E0 from the attributes, with nelec the total electron count of the walkers (used
 * by the Madelung term). */
int afq_h5_read_hamiltonian(const char* path, int nelec, double* hcore, double* L, double* E0);

/* /Wavefunction: kind (AFQ_H5_NOMSD / AFQ_H5_PHMSD), walker type of spin_type, orbitals N,
 * electrons per spin, determinants, and the number of PsiT_k groups. */
int afq_h5_wavefunction_info(const char* path, int* kind, int* walker_type, int* N, int* nup,
                             int* ndown, int* ndets, int* npsit);

/* Dense complex [rows][cols] of /Wavefunction/<kind>/PsiT_k (rows, cols must match the file). */
int afq_h5_read_psit(const char* path, int kind, int k, int rows, int cols, double* out);

/* ci_coeffs [ndets] (complex). */
int afq_h5_read_ci(const char* path, int kind, int ndets, double* ci);

/* PHMSD occupations: occa [ndets][nup], occb [ndets][ndown] (the first ndets rows). */
int afq_h5_read_occupations(const char* path, int ndets, int nup, int ndown, int* occa, int* occb);

/* ---- results file: the bins of the estimators' observables --------------------------------
 * The layout Estimators::write (Estimators.hpp:115-137), Measurements::write
 * (Measurements.hpp:69-80) and Accumulator::write (Accumulator.hpp:34-143) produce: the root
 * carries the integer attribute format_version = 1; observable `name` (a '/'-separated path
 * such as "Energy" or "BackPropEstimator/Steps=100/OneRDM") of stage `stage` ("Stage0", ...)
 * is the dataset Measurements/<stage>/<name>/bins of shape [bins][dims...], with a trailing
 * axis of 2 and the string attribute __complex__ = "1" for complex values, extendable along
 * the bin axis (chunked by one append, deflate level 1).
 * afq_h5_results_append creates the file when it does not exist and appends nbins bins
 * (data: nbins * prod(dims) values, complex interleaved) through a copy renamed over the file
 * at the end, as the reference does; the dims and the complex flag must match an existing
 * dataset. */
int afq_h5_results_append(const char* path, const char* stage, const char* name, int rank,
                          const long* dims, int is_complex, long nbins, const double* data);
/* Bin count, rank and dims (at most 8, without the complex axis) and complex flag of an
 * observable's bins. */
int afq_h5_results_shape(const char* path, const char* stage, const char* name, long* nbins,
                         int* rank, long* dims, int* is_complex);
/* All bins of an observable (nbins * prod(dims) values, complex interleaved). */
int afq_h5_results_read(const char* path, const char* stage, const char* name, double* out);
/* Observable names of a stage, '\n'-separated, into buf of size cap (the length needed,
 * including the terminating zero, in *len). */
int afq_h5_results_names(const char* path, const char* stage, char* buf, long cap, long* len);

#ifdef __cplusplus
}
#endif

#endif /* AFQ_H5_H */
