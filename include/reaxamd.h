/* C ABI of the MI355X ReaxFF engine (libreaxamd.so): the boundary a LAMMPS plugin (pair_style reaxff/metal, fix qeq/reaxff/metal)
 * or any other host binds. Plain C types only; every call returns 0 on success or a negative status, with the message available
 * from rxa_last_error(); nothing throws across this boundary. One context = one force field + one atom set + one HIP stream on one
 * device; a context is not thread-safe.
 *
 * Each entry point names the reference interface it replaces (yeatanir/Metal-Reax; paths relative to that repository).
 * The reference has no C ABI: its LAMMPS adapter calls the C++ engine directly (plugin/adapter/pair_reaxff_metal.cpp,
 * fix_qeq_reaxff_metal.cpp) and the device through reaxmetal::mtl::Context (include/reaxmetal/metal_backend.hpp:57-95).
 *
 * Atoms are the OWNED atoms of a periodic cell (the adapter reduces LAMMPS' ghosts to owner + lattice shift, as the reference's
 * host_atom_set does at plugin/adapter/pair_reaxff_metal.cpp:244-291); every periodic image is handled inside the engine, and the
 * returned gradient is already folded onto the owned atoms (ghost forces are zero).
 */
#ifndef REAXAMD_H
#define REAXAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rxa_context rxa_context;

/* status codes */
#define RXA_OK 0
#define RXA_EINVAL -1     /* bad argument / unsupported input */
#define RXA_EFFIELD -2    /* force field, control file or pair_style syntax error */
#define RXA_EDEVICE -3    /* HIP error */
#define RXA_ESTATE -4     /* call out of order (e.g. compute before atoms) */
#define RXA_ENOCONV -5    /* charge equilibration did not converge (strict) */

/* energy slots of rxa_compute: e_bond e_ov e_un e_lp e_ang e_pen e_coa e_hb e_tor e_con e_vdW e_ele e_pol
 * (reference include/reaxmetal/energy_terms.hpp:17-32; rxa_pvector maps them to LAMMPS' 14-slot pvector, energy_terms.cpp:48-63) */
#define RXA_NTERMS 13

/* ---- lifecycle ---------------------------------------------------------------------------------------------------
 * replaces: PairReaxFFMetal::metal_context() lazy creation (plugin/adapter/pair_reaxff_metal.cpp:370-374),
 *           mtl::Context::Context() / ~Context() (include/reaxmetal/metal_backend.hpp:59-60) */
int rxa_create(rxa_context** ctx, int device);
void rxa_destroy(rxa_context* ctx);
const char* rxa_last_error(const rxa_context* ctx);
/* ABI version of this header (major * 100 + minor) */
int rxa_abi_version(void);

/* ---- pair_style / pair_coeff --------------------------------------------------------------------------------------
 * rxa_pair_style replaces PairReaxFFMetal::settings -> parse_pair_style_args (plugin/adapter/pair_reaxff_metal.cpp:86-101,
 *   src/io/pair_settings.cpp:41-91): args[0] = control file name or "NULL", then keyword/value pairs (checkqeq, enobonds, lgvdw,
 *   shellcheck, backend cpu64|gpu (metal accepted), bonded cpu64|gpu, precision fp64|fp32|mixed, safezone, mincap, minhbonds,
 *   list/blocking, tabulate). control_text = contents of the control file (NULL when args[0] is "NULL").
 * rxa_pair_coeff replaces PairReaxFFMetal::coeff (pair_reaxff_metal.cpp:103-164) -> read_force_field_file
 *   (src/io/ffield_parser.cpp:91-460): run-time ffield reader; elements[t] is the element of LAMMPS type t+1 or "NULL".
 *   map_out (ntypes entries, may be NULL) receives the ffield index of each LAMMPS type (-1 for NULL).
 * rxa_extract replaces PairReaxFFMetal::extract("chi"|"eta"|"gamma") (pair_reaxff_metal.cpp:212-227): out[0..ntypes], index = LAMMPS
 *   type, out[0] unused; eta is 2 x the file value, as LAMMPS stores it. */
int rxa_pair_style(rxa_context* ctx, int nargs, const char* const* args, const char* control_text);
int rxa_pair_coeff(rxa_context* ctx, const char* ffield_text, int ntypes, const char* const* elements, int* map_out);
int rxa_extract(rxa_context* ctx, const char* name, double* out);
/* out[0] nonb_cut, out[1] bond_cut, out[2] hbond_cut, out[3] required ghost shell max(nonb, hbond, 2 bond)
 * (replaces PairReaxFFMetal::cutoffs, pair_reaxff_metal.cpp:230-236, and NeighborCutoffs::required_shell, src/neighbor/neighbor.cpp:15) */
int rxa_cutoffs(rxa_context* ctx, double* out);
/* canonical table dump (same text as the reference's src/io/table_dump.cpp:20-84); returns the length, copies up to cap-1 chars */
long rxa_table_dump(rxa_context* ctx, int portable, char* buf, long cap);

/* ---- atoms (replaces host_box / host_atom_set / make_device_list_input, pair_reaxff_metal.cpp:238-291, src/neighbor/neighbor.cpp:329-354)
 * LAMMPS box: lo, hi, tilt (xy, xz, yz), periodic flags. types are LAMMPS types (1..ntypes), tags the atom ids. */
int rxa_set_box(rxa_context* ctx, const double lo[3], const double hi[3], const double tilt[3], const int periodic[3]);
int rxa_set_atoms(rxa_context* ctx, int64_t n, const int* types, const int64_t* tags, const double* x);
/* LAMMPS' ghost-native arrays as they are (nlocal owned atoms, then nghost ghosts; types/tags/x of length nlocal + nghost),
 * replacing host_atom_set (pair_reaxff_metal.cpp:244-291): every ghost must be a lattice image of the owned atom with its id
 * and type, shifted along periodic directions only (checked to 1e-8 of the box scale, as AtomSet::validate); then the owned
 * atoms are loaded as by rxa_set_atoms (the engine builds every image itself, and grad/peratom cover the owned atoms only:
 * ghost forces stay zero). nranks = comm->nprocs: more than one rank returns RXA_EINVAL, since a sub-domain's ghosts are owned
 * by other ranks and this engine evaluates whole periodic cells (the reference's multi-rank 'distributed' atom set). */
int rxa_set_atoms_ghosted(rxa_context* ctx, int64_t nlocal, int64_t nghost, const int* types, const int64_t* tags, const double* x,
                          int nranks);
/* the same checks without a context (no device needed): RXA_OK, or the error code with its message copied to err */
int rxa_check_ghosts(const double lo[3], const double hi[3], const double tilt[3], const int periodic[3], int64_t nlocal, int64_t nghost,
                     const int* types, const int64_t* tags, const double* x, int nranks, char* err, long errcap);
int rxa_set_positions(rxa_context* ctx, const double* x); /* 3n, owned atoms, same order */
int rxa_set_charges(rxa_context* ctx, const double* q);   /* n */
int rxa_get_charges(rxa_context* ctx, double* q);         /* n */

/* ---- forces (replaces PairReaxFFMetal::compute, pair_reaxff_metal.cpp:412-524, i.e. mtl::Context::bonded + nonbonded,
 * metal_backend.hpp:78-80, and compute_bonded_core / compute_nonbonded_core, include/reaxmetal/bonded.hpp:56, nonbonded.hpp:35)
 * energies[RXA_NTERMS] (may be NULL), grad[3n] = dE/dx of the owned atoms (the adapter SUBTRACTS it from atom->f, as the reference
 * does at pair_reaxff_metal.cpp:501-504; may be NULL), virial[6] = xx yy zz xy xz yz of all terms (may be NULL). */
#define RXA_COMPUTE_REBUILD 1   /* force a neighbour-list rebuild */
#define RXA_COMPUTE_PERATOM 2   /* also tally per-atom energies and virials (read them with rxa_peratom) */
#define RXA_COMPUTE_CENSUS 4    /* also record the decision census (read it with rxa_census) */
int rxa_compute(rxa_context* ctx, int flags, double* energies, double* grad, double* virial);
/* per-atom energy eatom[n] and virial vatom[6n] (xx yy zz xy xz yz per atom) of the last rxa_compute made with
 * RXA_COMPUTE_PERATOM, owned atoms in the caller's order, every image folded onto its owner (either pointer may be NULL).
 * Replaces the per-atom tallies of PairReaxFFMetal::compute with eflag_atom / vflag_atom (pair_reaxff_metal.cpp:412-524), which
 * the reference takes from its CPU engine (BondedOptions::per_atom / NonbondedOptions::per_atom, include/reaxmetal/bonded.hpp:28,
 * nonbonded.hpp:22). Energies follow its attribution (bonded.cpp:212-704, nonbonded.cpp:55-76): bond, triple-bond and C2 terms
 * half to each atom of the bond, lone pair and over/under to the atom, valence to the centre, torsion half to each atom of the
 * central bond, H-bond a third to each atom of the triple, van der Waals and Coulomb half to each atom of the pair,
 * polarisation to the atom. The virial follows the reference's split (bonded.cpp:462-466, 636-647, 703-710, 737-766):
 * non-bonded pairs half to each atom, each valence angle's virial a third to each of its atoms, each torsion's a quarter to each
 * of four, each H-bond triple's a third to each of three, and the bond-order derivative pass per bond as the reference tallies
 * it; the per-atom virials add up to the virial of rxa_compute. */
int rxa_peratom(rxa_context* ctx, double* eatom, double* vatom);
/* the 13 terms -> LAMMPS' 14-slot pvector (eb ea elp emol ev epen ecoa ehb et eco ew ep efi eqeq), energy_terms.cpp:48-63 */
void rxa_pvector(const double* energies, double* pvector14);
/* per owned atom decision census, 7 * n ints [k * n + atom] in the caller's order (replaces BondedOptions::census / DecisionCensus,
 * include/reaxmetal/census.hpp:12-20; the reference fills it in src/cpu/bonded.cpp and nonbonded.cpp):
 *   0 bonds (the bond-order cut-off), 1 angle_sets: bond pairs of the valence centre through the three-body gates (BO > thb_cut
 *   each, product > thb_cutsq), 2 torsions: chains through the four-body gates with parameters, at the central bond's counted
 *   end, 3 hbonds: triples evaluated at the donor hydrogen, 4 nonbonded: pairs within nonb_cut at the counted end, 5 sbo_region:
 *   0 SBO <= 0, 1 (0, 1], 2 (1, 2), 3 >= 2, 6 lp_trunc: (int)(Delta_e / 2).
 * The kernels that take the decisions record them (in each mode's own precision). Reads the census of the last rxa_compute made
 * with RXA_COMPUTE_CENSUS; otherwise first runs such a force call at the current positions and charges. */
int rxa_census(rxa_context* ctx, int* out);
/* bond table of the owned atoms: start[n+1], nbr[], bo[] (nbr = owned index of the partner), total_bo[n], nlp[n]; returns the
 * number of bonds; call with nbr == NULL to size (replaces BondResult::table for fix reaxff/bonds|species, bonded.hpp:29-35,
 * publish_bonds at pair_reaxff_metal.cpp:330-368) */
long rxa_bond_table(rxa_context* ctx, int32_t* start, int32_t* nbr, double* bo, double* total_bo, double* nlp);

/* ---- charge equilibration (fix qeq/reaxff/metal; replaces mtl::Context::qeq_setup / qeq_matvec / qeq_solve, metal_backend.hpp:81-86,
 * FixQEqReaxFFMetal::compute_H / sparse_matvec / CG / calculate_Q, plugin/adapter/fix_qeq_reaxff_metal.cpp:65-205)
 * chi, eta, gamma indexed by LAMMPS type (index 0 unused), as the fix holds them (eta = 2 x ffield value). */
int rxa_qeq_setup(rxa_context* ctx, double swa, double swb, const double* chi, const double* eta, const double* gamma);
/* y = (diag(eta) + H) x over the owned atoms (the fix's sparse_matvec; images folded) */
int rxa_qeq_matvec(rxa_context* ctx, const double* x, double* y);
/* The charge-matrix product y = (diag(eta) + H) x of rxa_qeq_matvec on DEVICE vectors in the caller's order (n doubles each),
 * allocating nothing: no counterpart in the reference; for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Its staging vectors live in a file-scope device array of the library
 * (2 x 262144 doubles, 4 MiB per device, held until the process ends): more than 262144 atoms returns RXA_EINVAL. The matrix
 * of the current positions must already be built (rxa_qeq_matvec, rxa_qeq_solve, rxa_qeq_charges or an MD step), else
 * RXA_ESTATE. Work is queued on the context's main stream (rxa_stream) and not waited for. Every context on one device shares
 * that device's array, so calls that overlap on one GPU race. */
int rxa_qeq_matvec_noalloc(rxa_context* ctx, const double* x_dev, double* y_dev);
/* both systems at once (b, x0, x: 2n = system s then system t); iters[2] = the stock loop index at exit, rel[2] = sqrt(r.M^-1 r)/|b|,
 * status[2]: 1 converged, 3 iteration limit (the reference's CgOutcome, metal_backend.hpp:84-85) */
int rxa_qeq_solve(rxa_context* ctx, const double* b, const double* x0, int imax, double tol, double* x, int* iters, double* rel, int* status);
/* the whole per-step charge update of the stock fix: history extrapolation, both solves, q = s - (sum s / sum t) t, history shift;
 * the charges stay on the device for rxa_compute and are copied to q_out (may be NULL). strict != 0 turns a non-converged solve
 * into RXA_ENOCONV (the fix's 'strict' keyword). */
int rxa_qeq_charges(rxa_context* ctx, int imax, double tol, int strict, double* q_out, int* iters, double* rel);
int rxa_qeq_reset_history(rxa_context* ctx);
/* FP64 iterative refinement of the charge solve (replaces the fix's 'resident' keyword, plugin/adapter/fix_qeq_reaxff_metal.cpp:38,
 * 67-80, i.e. mtl::Context::qeq_solve's double-single refinement, src/metal/metal_backend.mm:435-530): the CG multiplies with the
 * FP32 matrix values, while its initial residual and a check of the true residual after the iterations stop use the matrix to
 * about 48 bits (FP32 value + FP32 remainder); a solve whose true residual fails the stock condition restarts from it, within
 * one imax budget. The charges then solve the FP64 matrix to the requested tolerance. mode -1: the precision's default (on in
 * mixed mode, off in FP32 mode), 0 off, 1 on; FP64 mode keeps FP64 values and is not affected. rxa_qeq_solve / rxa_qeq_charges
 * then report the total iterations (+1) and the true relative residual of the last check. */
int rxa_qeq_refine(rxa_context* ctx, int mode);

/* ---- standalone MD loop (bench.py and the serving harness; LAMMPS integrates in plugin mode)
 * masses by LAMMPS type (index 0 unused); NVT Nose-Hoover chain (3) at t_target with damping t_damp (fs), or NVE when t_damp <= 0;
 * dt in fs; skin in Angstrom; qeq_every 0 keeps charges fixed. */
int rxa_md_setup(rxa_context* ctx, const double* v, const double* mass_by_type, double dt, double t_target, double t_damp, double skin,
                 int qeq_every, int qeq_imax, double qeq_tol);
/* runs `steps` steps; thermo (may be NULL) receives: potential energy, kinetic energy, temperature, neighbour rebuilds, mean QEq
 * iterations (s + t)/2, last iters_s, last iters_t, max displacement since the last rebuild */
int rxa_md_run(rxa_context* ctx, int steps, double* thermo);

/* ---- several independent systems in one context (no counterpart in the reference, which runs one system per context:
 * capabilities.cpp:112; for batched and served MD, where one launch of each kernel then covers every system)
 * nsys systems, counts[nsys] atoms each, stored one after another in types / tags / x (3 per atom); boxes[9 nsys] = lo[3], hi[3],
 * tilt[3] per system; periodic[3 nsys]. Atom ids are per system. Replaces rxa_set_box + rxa_set_atoms; every other call then
 * acts on all systems at once (positions, charges, gradient: all atoms in this order; energies and thermo: totals). Each
 * system keeps its own neighbour rebuilds, charge solve (own scalars and stopping), thermostat and energy sums, so its results
 * do not depend on the systems packed with it. */
int rxa_set_systems(rxa_context* ctx, int nsys, const int64_t* counts, const double* boxes, const int* periodic, const int* types,
                    const int64_t* tags, const double* x);
/* per-system results of the last force call / MD run: energies[RXA_NTERMS nsys], virial[6 nsys], thermo[4 nsys] = kinetic energy,
 * temperature, stock loop index of the s and t solves of the last charge solve (each may be NULL) */
int rxa_system_results(rxa_context* ctx, double* energies, double* virial, double* thermo);
/* per-system thermostat target (K) and damping (fs; <= 0: NVE) after rxa_md_setup (either may be NULL) */
int rxa_md_system_targets(rxa_context* ctx, const double* t_target, const double* t_damp);
int rxa_get_state(rxa_context* ctx, double* x, double* v, double* q);
/* ---- continuing a run in a new pack (no counterpart in the reference; for the continuous-batching server, which re-packs
 * the running systems with newly admitted ones at chunk boundaries)
 * rxa_md_state: the state a run continues from beyond x, v and q: the charge-solve history of s and t (s_hist[4n], t_hist[4n],
 * [k n + atom], k = 0 newest, the caller's order) and the thermostat chain of each system (nh[32 nsys]; 32 for one system).
 * rxa_md_resume: rxa_md_setup for a context whose systems (rxa_set_systems) continue runs where keep[k] != 0: those take the
 * given charges q[n], history and chain (values of the other systems are ignored) instead of a first charge solve, so their
 * next step is the one their previous context would have taken (the neighbour lists are rebuilt, which changes only the order
 * of sums); the other systems start as rxa_md_setup starts them. Charge history is required when qeq_every > 0. */
/* ---- autotuner (no counterpart in the reference): times launch configurations on the running MD and keeps the fastest of
 * each in turn: lanes per atom of the slot-parallel per-atom kernels (1, 2, 4, 8), the angle kernels' group size (the largest
 * bond count's power of two, or twice it), the bonded chain on a second stream beside the charge solve or not, and the
 * non-bonded pass over the charge-matrix rows or over the skin list. Every configuration gives the same results. Each value
 * is timed twice, the values in order and then in reverse, and scored by the sum of its two trials (20 trials); each trial
 * advances the run by 2 + 2 x steps MD steps (after rxa_md_setup). choice[5] = lanes, group multiplier, overlap, rows, trials;
 * trial_ms (may be NULL) receives up to trial_cap trials' ms per step in the order tried. */
int rxa_autotune(rxa_context* ctx, int steps, int64_t* choice, double* trial_ms, int trial_cap);
/* the launch configuration the autotuner chooses, set and read directly: [4] = lanes per atom (1, 2, 4, 8), group multiplier
 * (1, 2), overlap (0, 1), non-bonded over the rows (0, 1); either pointer may be NULL */
int rxa_launch_config(rxa_context* ctx, const int* set, int* get);
int rxa_md_state(rxa_context* ctx, double* s_hist, double* t_hist, double* nh);
int rxa_md_resume(rxa_context* ctx, const double* v, const double* mass_by_type, double dt, double t_target, double t_damp, double skin,
                  int qeq_every, int qeq_imax, double qeq_tol, const int* keep, const double* q, const double* s_hist,
                  const double* t_hist, const double* nh);

/* capacities and counters: out[0] n, [1] skin-list capacity per atom, [2] bond-slot capacity, [3] H-bond candidate capacity,
 * [4] skin-list entries, [5] directed bond slots in use, [6] acceptors within hbond_cut summed over the donors (last evaluation),
 * [7] charge-matrix entries (full rows), [8] neighbour rebuilds, [9] precision (0 fp64, 1 mixed, 2 fp32), [10] bond candidates
 * (r <= bond_cut + skin, directed), [11] H-bond candidates (donor and acceptor lists, r <= hbond_cut + skin) */
int rxa_stats(rxa_context* ctx, int64_t* out);
/* charge-solve counters since the context's solve buffers were sized: out[0] solves (rxa_qeq_charges / MD steps), out[1]
 * refinement checks whose two-float product ran (the others were settled by the bound), out[2] restarts from the true residual,
 * out[3] batches of the single-system solve (each one host read-back of its flags) */
int rxa_qeq_stats(rxa_context* ctx, int64_t* out);

/* ---- context-free helpers (no device is touched; usable on a host without a GPU) ----------------------------------
 * rxa_ffield_table_dump: run-time reader + canonical dump of a force field text (0 on success, length in *len); err receives the
 * reader's message on failure. rxa_check_pair_style: parses pair_style arguments; summary receives "key=value" lines of the
 * effective settings (precision, backend, bonded, cutoffs, notices). Both replace the reference's standalone parser entry points
 * (src/io/ffield_parser.cpp read_force_field_file + table_dump.cpp canonical_table_dump; src/io/pair_settings.cpp:41-91). */
int rxa_ffield_table_dump(const char* ffield_text, int lgvdw, int portable, char* buf, long cap, long* len, char* err, long errcap);
int rxa_check_pair_style(int nargs, const char* const* args, const char* control_text, char* summary, long cap);

/* ---- device timing of the engine's own stream (HIP events around each phase; off by default)
 * phases: 0 neighbour rebuild, 1 bond lists, 2 bond-order correction, 3 bonded terms, 4 non-bonded, 5 bond-derivative gathers +
 * energy reduction, 6 charge matrix, 7 charge SpMV (one launch per CG iteration), 8 other CG work, 9 integrator.
 * rxa_profile(mask): 0 off, 1 every phase, otherwise a bit mask of the phases to time (bit k = phase k, e.g. 1 << 7 times only
 * the SpMV launches); it also clears the totals. rxa_profile_read fills ms[10] (summed milliseconds) and counts[10] (timed
 * launches). */
#define RXA_NPHASES 10
int rxa_profile(rxa_context* ctx, int mask);
int rxa_profile_read(rxa_context* ctx, double* ms, int64_t* counts);
/* time only every `every`-th launch sequence of each timed phase (1 = all; clears the totals): fewer events in a timed loop.
 * rxa_profile_read then reports every launch sequence in counts and the timed mean times that count in ms. No reference
 * counterpart (measurement of this engine). */
int rxa_profile_every(rxa_context* ctx, int every);

/* ---- device-resident inputs (bench / serving): pointers to device memory owned by the caller (HIP, same device) */
int rxa_set_positions_device(rxa_context* ctx, const double* dx);
int rxa_stream(rxa_context* ctx, void** hip_stream_out);
/* profiling aid (no counterpart in the reference): `reps` launches of a plain copy kernel over `bytes` (a multiple of 16) between
 * two scratch buffers on the context's stream, so memory counters can be calibrated against a kernel of known traffic */
int rxa_stream_copy(rxa_context* ctx, int64_t bytes, int reps);

#ifdef __cplusplus
}
#endif
#endif
