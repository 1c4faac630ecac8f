/*
 * hotspice_hip.h -- C-ABI of libhotspice_hip.so, the gfx950 (MI355X) implementation of the
 * Hotspice Monte Carlo stepping path.
 *
 * The library replaces the array expressions that Hotspice's stepping path evaluates with
 * NumPy/CuPy (the reference has no native code). Every entry point below names the
 * reference function(s) it stands in for. The host-side Python package hotspice_amd binds
 * these with ctypes; its classes keep the reference's Magnets / Energy API.
 *
 * Conventions
 *  - All array pointers are DEVICE pointers (HBM), row-major (ny, nx) grids indexed [y][x],
 *    flat index c = y*nx + x. Memory is owned by the caller (PyTorch tensors in hotspice_amd).
 *  - Every function returns 0 on success and a negative code on failure; hs_last_error()
 *    returns a message for the calling thread. Launches are asynchronous on `stream`
 *    (a hipStream_t passed as void*); only functions documented as blocking synchronise.
 *  - Energies are held as ONE combined local field per magnet (the sum over the Energy
 *    objects the reference keeps separately, energies.py:22-83) in FP32 with a Kahan
 *    compensation array: value = E - Ec. Likewise for the perpendicular field E_perp.
 *  - Random numbers are Philox4x32-10 streams: key = seed, counter = (index, step, tag);
 *    see hs_philox_uniform53.
 */
#ifndef HOTSPICE_HIP_H
#define HOTSPICE_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HS_MAX_UC 64  /* max unit-cell sites ux*uy */
#define HS_MAX_NN 49  /* max entries of the exchange nearest-neighbour mask (7x7) */

/* Cell flag bits (uint8 per cell). */
#define HS_FLAG_OCC 1u   /* occupied now (Magnets.occupation, used by selection core.py:485) */
#define HS_FLAG_NEEL 2u  /* finite Magnets._occupation_inf (core.py:95-96): Neel candidate */

/* Error codes. */
#define HS_OK 0
#define HS_ERR_ARG -1
#define HS_ERR_HIP -2
#define HS_ERR_STATE -3

/*
 * One spin-ice instance: geometry, parameters and device pointers. Field order is part of
 * the ABI (mirrored by hotspice_amd/_lib.py). Layout: 8-byte members first.
 */
typedef struct hs_instance {
  /* device pointers */
  int8_t *m;            /* (ny,nx) spins -1/0/+1 */
  uint8_t *flags;       /* (ny,nx) HS_FLAG_* */
  float *E, *Ec;        /* combined local field and its Kahan compensation */
  float *Ep, *Epc;      /* combined perpendicular field (without m_perp_factor) */
  const float *EB;      /* (ny,nx) E_B */
  const float *beta;    /* (ny,nx) 1/(kB T) */
  const float *mu;      /* (ny,nx) pair weight of each magnet, FP32 (stepping): its moment for
                           DipolarEnergy, its squared charge (moment/d)^2 for DiMonopolarEnergy */
  const float *mu_src;  /* (ny,nx) weight of a switching magnet in the pair terms it sends, FP32:
                           == mu for DipolarEnergy, ones for DiMonopolarEnergy */
  const double *mu64;   /* (ny,nx) moment, FP64 (Zeeman, totals) */
  const double *pw64;   /* (ny,nx) pair weight, FP64 (recompute); == mu64 for DipolarEnergy */
  const float *K32;     /* [n_sub][2kwy+1][2kwx+1] dipolar kernel, FP32 (stepping) */
  const float *Kpo32;   /* same shape, K_perpother (stepping E_perp) */
  const double *K64;    /* [n_sub][2kwy+1][2kwx+1] dipolar kernel, FP64 (recompute) */
  const double *Kps64;  /* same shape, K_perpself (recompute E_perp) */
  const float *zee_bx;  /* optional (ny,nx) field x-component (OOP: magnitude); NULL = uniform */
  const float *zee_by;  /* optional (ny,nx) field y-component; NULL = uniform */
  const int32_t *occ_list; /* [n_occ] flat indices of occupied cells, row-major (core.py:97-98) */
  const double *beta64c;  /* (ny,nx) 1/(kB T) in FP64: Neel stepping with uniform_beta = 0 needs it
                             (the per-cell waiting times); NULL otherwise */
  double *E64;          /* (ny,nx) combined local field in FP64, or NULL. When set, the instance holds
                           its field here instead of in E/Ec (which are then unused): the recompute
                           writes it, readouts and Zeeman steps use it, and Metropolis steps on grids
                           of at most 80x80 cells update it with FP64 increments from K64 and mu64
                           (DipolarEnergy / ExchangeEnergy / ZeemanEnergy, no E_perp). Every other
                           stepping entry point refuses such an instance with HS_ERR_ARG. */
  double *C64;          /* (2, ny, nx) FP64 shadow of the raw dipolar correlations of the last full
                           recompute, sum_j K_n(i)[j - i] (m w)_j for E and (K_perpself) for E_perp,
                           or NULL; a recompute of both components (which = 3, or 1 without E_perp)
                           stores it, hs_refresh_fields updates it */
  int8_t *m_ref;        /* (ny,nx) the spins C64 was computed for, or NULL (set with C64) */
  uint64_t seed;        /* Philox key */
  double zee_Bx, zee_By; /* uniform field components when zee_bx/zee_by are NULL */
  double dip_pf;        /* DipolarEnergy.prefactor (0 = no dipolar energy) */
  double J;             /* ExchangeEnergy.J (0 = no exchange energy) */
  double m_perp_factor; /* Magnets.m_perp_factor */
  double beta64;        /* 1/(kB T) in FP64 when uniform_beta: Neel waiting times and
                           METROPOLIS_TIME step times exp(b*beta) need it at ~1e-9 relative,
                           since b*beta reaches a few hundred */
  /* 4-byte members */
  int32_t nx, ny, ux, uy;
  int32_t n_sub;        /* number of dipolar kernels (occupied unit-cell sites) */
  int32_t kwx, kwy;     /* stored kernel half widths (nx-1, ny-1 for the full kernel) */
  int32_t pbc;          /* periodic boundaries */
  int32_t in_plane;
  int32_t use_perp;     /* Magnets.USE_PERP_ENERGY (core.py:273-275) */
  int32_t barrier;      /* 0 = 'simple', 1 = 'parabolic' (core.py:641-648) */
  int32_t nn_w, nn_h;   /* exchange NN mask width/height (odd) */
  int32_t n_occ;        /* Magnets.n */
  int32_t has_zeeman;
  int32_t uniform_beta; /* 1 when T is uniform over the grid */
  int32_t pair_src_unit; /* 0: a pair term is prefactor*K*(m w)_source*(m w)_receiver (DipolarEnergy);
                            1: prefactor*K*m_source*(m w)_receiver (DiMonopolarEnergy), w = mu/pw64;
                            the stepping kernels read the source weight from mu_src */
  int32_t sub_of[HS_MAX_UC];    /* unit-cell site (y%uy)*ux + x%ux -> kernel index or -1 */
  double ox_uc[HS_MAX_UC];      /* in-plane orientation per unit-cell site */
  double oy_uc[HS_MAX_UC];
  float nn_mask[HS_MAX_NN];     /* exchange NN mask, [nn_h][nn_w] */
} hs_instance_t;

/* Geometry needed to build dipolar kernels (DipolarEnergy._initialize, energies.py:176-288). */
typedef struct hs_kernel_geom {
  const double *rrx;  /* [n_sub][2nx-1] x-offsets (m) from the centre magnet, host-computed cumsums */
  const double *rry;  /* [n_sub][2ny-1] y-offsets (m) */
  double I;           /* (major^2 + minor^2)/16 finite-size term */
  double cutoff;      /* cutoff_range (m), +inf for none */
  double decay;       /* decay_exponent (-3 by default) */
  int32_t nx, ny, ux, uy;
  int32_t n_sub;
  int32_t kwx, kwy;   /* half widths of the stored window */
  int32_t pbc, in_plane;
  int32_t site_x[HS_MAX_UC], site_y[HS_MAX_UC]; /* unit-cell site of kernel n */
  int32_t shift_x[HS_MAX_UC], shift_y[HS_MAX_UC]; /* tiled-matrix slice starts (energies.py:205-206) */
  double uc_ox[HS_MAX_UC], uc_oy[HS_MAX_UC]; /* orientation per site (IP) */
  double uc_o[HS_MAX_UC];                    /* occupation per site (OOP) */
} hs_kernel_geom_t;

/* Dumbbell geometry of DiMonopolarEnergy (energies.py:406-463), per unit-cell site s = y*ux + x. */
typedef struct hs_monopole_geom {
  double ca[HS_MAX_UC], sa[HS_MAX_UC];  /* cos, sin of the site's original angle */
  double cp[HS_MAX_UC], sp[HS_MAX_UC];  /* cos, sin of (that angle + pi/2) */
  double h[HS_MAX_UC];                  /* d/2: half the distance between the site's two charges */
  double h_small;                       /* small_d/2, for a magnet turned perpendicular */
} hs_monopole_geom_t;

/* Run control for a stepping call (Magnets.progress, core.py:736-760). */
typedef struct hs_run {
  double t_max;         /* progress t_max (may be +inf) */
  double mc_max;        /* progress MCsteps_max (may be +inf) */
  double attempt_freq;  /* f0 (core.py:659, 679) */
  int64_t max_steps;    /* step budget of this call */
  int32_t log;          /* 1: write per-step log arrays */
  int32_t nblocks;      /* grid size for grid-wide kernels (0 = automatic); Metropolis: workgroups
                           of the update launch over the step's tile worklist (0 = 2048) */
  int32_t begin;        /* 1: this call starts a progress() call (t0, attempts0 := t, attempts) */
  int32_t first_free;   /* 1: the call's first step runs whatever t and MCsteps are, as
                           Magnets.update steps unconditionally (core.py:659-718);
                           2: every step does (a sequence of update() calls, each Neel step
                           compared with t_max itself, core.py:702); 0: progress() */
} hs_run_t;

/* Device-resident run state; the library keeps two copies (double buffer). */
typedef struct hs_state {
  double t;             /* Magnets.t */
  double t0;            /* t at the start of the progress call */
  int64_t attempts;     /* Magnets.attempted_switches */
  int64_t attempts0;
  int64_t switches;     /* Magnets.switches */
  int64_t step;         /* update() calls so far: the Philox step counter */
  int64_t steps_done;   /* steps done by this call */
  int32_t done;         /* 1 once the stop predicate fired */
  int32_t err;          /* nonzero on a device-side error */
} hs_state_t;

/* Metropolis selection geometry for one r (Magnets._select_grid, core.py:496-558, or the
 * Poisson-disc scheme of Magnets._select_Poisson, core.py:560-570), or the caller's own samples
 * (scheme 3: update(idx=...), core.py:660-661): pat_cells holds nsc flat cell indices y*nx+x in
 * the caller's order, every one a magnet (a step whose list holds anything else ends with
 * err = 4 before it changes anything). */
typedef struct hs_select {
  const int32_t *pat_off;   /* [uy*ux+1] offsets into pat_cells for supercell pattern (dy,dx) */
  const int32_t *pat_cells; /* occupied (oy*Rx+ox) cells of occupation[dy:dy+Ry, dx:dx+Rx], row-major */
  int32_t Rx, Ry;           /* supercell size in cells */
  int32_t snx, sny;         /* supercells along x/y: n//R + 1 */
  int32_t ncols, nsc;       /* chosen (every other) supercells per row and in total; Poisson: background cells */
  int32_t scheme;           /* 0 = 'grid', 1 = 'single', 2 = 'Poisson', 3 = explicit sample list */
  int32_t cutoff;           /* SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF */
  int32_t reduced_k;        /* REDUCED_KERNEL_SIZE (0 = full kernel) */
  int32_t metropolis_time;  /* SimParams.METROPOLIS_TIME */
  double r_x, r_y;          /* Poisson: minimal sample distance r in cells along x (r/dx) and y (r/dy);
                               the background grid is snx x sny cells of (r_x, r_y)/sqrt(2) */
} hs_select_t;

/*
 * One spin-ice instance of a batched (persistent, multi-instance) launch. Every instance gets
 * a team of workgroups that steps it independently of the other teams; a team of one
 * workgroup synchronises with barriers only, larger teams through a bounded-spin atomic
 * barrier of their own.
 */
typedef struct hs_team {
  const hs_instance_t *inst;  /* HOST pointer to the instance descriptor (copied by the call) */
  const hs_select_t *sel;     /* HOST pointer, Metropolis teams only */
  hs_run_t run;
  hs_state_t *state;          /* device state (copy 0 is read and written) */
  int32_t *log_idx;           /* optional device logs: Neel winner (-1 = no switch) / Metropolis counts */
  double *log_dt;             /* Neel: time added per step */
  int32_t *log_cells;         /* Metropolis: accepted cells, capacity sel->nsc per step */
  int32_t nblocks;            /* workgroups of this team (0 = automatic) */
  int32_t pad;
} hs_team_t;

/* ---- entry points ----------------------------------------------------------------------- */

/* Library version (major*10000 + minor*100 + patch). */
int hs_version(void);

/* Message of the last error on this thread ("" if none). */
const char *hs_last_error(void);

/*
 * Layout of the ABI structs for binding self-checks: writes {sizeof(hs_instance_t),
 * offsetof seed, nx, sub_of, nn_mask; sizeof(hs_kernel_geom_t), offsetof nx, uc_o;
 * sizeof(hs_run_t), offsetof max_steps, begin; sizeof(hs_state_t), offsetof done;
 * sizeof(hs_select_t), offsetof metropolis_time; sizeof(hs_team_t), offsetof run, nblocks}
 * to out[0..17] when n >= 18. Returns the count (18). Touches no device. No counterpart in
 * the reference (binding infrastructure).
 */
int hs_abi_layout(int64_t *out, int32_t n);

/* Number of HIP devices and the name/arch of device 0 (blocking). */
int hs_device_info(int32_t *count, char *name, int32_t name_len);

/*
 * Philox4x32-10 53-bit uniforms out[i] for index = index0 + i under (seed, step, tag);
 * open_zero=0 gives [0,1), 1 gives (0,1]. Exposes the exact streams the stepping kernels
 * draw (replacing mm.rng / numpy.random at core.py:83, 493, 527-538, 671, 698) for replay.
 */
int hs_philox_uniform53(uint64_t seed, uint64_t step, uint32_t tag, int64_t index0, int64_t n,
                        int32_t open_zero, double *out, void *stream);

/*
 * Dipolar kernels K, K_perpself, K_perpother of every occupied unit-cell site, FP64, plus FP32
 * copies of K and K_perpother. Replaces DipolarEnergy._initialize (energies.py:176-288),
 * including the cutoff (219), PBC folding (263-277) and the mu0/4pi factor (279-281).
 * Outputs have shape [n_sub][2kwy+1][2kwx+1]; Kps64/Kpo64 may be NULL for OOP lattices.
 */
int hs_build_dipolar_kernels(const hs_kernel_geom_t *g, double *K64, double *Kps64, double *Kpo64,
                             float *K32, float *Kpo32, void *stream);

/*
 * The same three kernels for DiMonopolarEnergy (energies.py:406-641): every magnet is a pair of
 * opposite charges d apart along its angle, and an entry is the sum of the four charge-charge
 * 1/r terms (like charges +, unlike -) between the centre magnet and the magnet at that offset,
 * times the target site's occupation. K takes the squared distances in FP32 as the reference
 * does (energies.py:508-525); K_perpself turns the centre magnet by +90 degrees with charges
 * small_d apart, K_perpother every other magnet. PBC folding and mu0/4pi as for the dipolar
 * kernels. The window is always the full one (kwx = nx-1, kwy = ny-1). Here g->rrx[n][C] and
 * g->rry[n][R] hold ABSOLUTE positions of the columns/rows of the reference's enlarged grid seen
 * from kernel n (energies.py:416-428): the centre magnet sits at C = nx-1, R = ny-1; g->uc_o is
 * the occupation per unit-cell site, g->I / cutoff / decay / in_plane are unused. mg is a DEVICE
 * pointer (the per-site tables are read by every thread; passed by value they would exceed the
 * kernel-argument space together with g).
 */
int hs_build_dimonopolar_kernels(const hs_kernel_geom_t *g, const hs_monopole_geom_t *mg, double *K64,
                                 double *Kps64, double *Kpo64, float *K32, float *Kpo32, void *stream);

/*
 * FP64 full recompute of the combined field (and E_perp when which&2), written as FP32 +
 * compensation. Replaces Energy.update() of DipolarEnergy (energies.py:290-308), ZeemanEnergy
 * (energies.py:116-118) and ExchangeEnergy (energies.py:771-781) as called by
 * Magnets.update_energy() with index=None (core.py:247-249). which: 1 = E, 2 = E_perp, 3 = both.
 * The dipolar sum runs per (target, source) sublattice pair as a direct correlation on the
 * unit-cell grids, split over groups of source tiles; scratch holds the per-group partial
 * sums (hs_recompute_scratch_bytes bytes).
 */
int hs_recompute_fields(const hs_instance_t *inst, int32_t which, void *scratch, int64_t scratch_bytes,
                        void *stream);

/* Scratch bytes hs_recompute_fields needs for this instance. */
int64_t hs_recompute_scratch_bytes(const hs_instance_t *inst);

/*
 * The FP64 recompute from the shadow (inst->C64, inst->m_ref, stored by the last full
 * hs_recompute_fields of both components): the cells whose spin differs from m_ref are listed in
 * cell order (per-row counts and an ordered compaction), their change (m - m_ref) w is added to the shadow
 * over their windows in that order, m_ref takes the spins, and the fields are finished from the
 * shadow as hs_recompute_fields finishes them (Zeeman and exchange closed forms included). Equal
 * to a full recompute up to FP64 rounding, at a cost of (changed cells x window) instead of
 * (cells x window): DipolarEnergy.update energies.py:290-308 by linearity in the sources. The
 * caller keeps the shadow valid: after any change of kernels, weights or occupation, a full
 * recompute comes first. scratch: hs_refresh_scratch_bytes bytes. Non-blocking.
 */
int hs_refresh_fields(const hs_instance_t *inst, int32_t which, void *scratch, int64_t scratch_bytes, void *stream);

/* Scratch bytes hs_refresh_fields needs for this instance (the ordered list of changed cells). */
int64_t hs_refresh_scratch_bytes(const hs_instance_t *inst);

/*
 * hs_recompute_fields for n instances at once (one launch for all dipolar correlations, one for
 * all finishing passes): insts[i] (HOST pointers to the descriptors), which[i] as above,
 * scratch[i] / scratch_bytes[i] the recompute scratch of instance i. ws: device scratch of
 * hs_recompute_many_ws_bytes bytes for the descriptors. Same results as n single calls.
 * Asynchronous: the descriptors are copied through pinned staging and the host arrays may be
 * released when the call returns.
 */
int hs_recompute_fields_many(const hs_instance_t *const *insts, const int32_t *which, void *const *scratch,
                             const int64_t *scratch_bytes, int32_t n, void *ws, int64_t ws_bytes, void *stream);
int64_t hs_recompute_many_ws_bytes(const hs_instance_t *const *insts, const int32_t *which, int32_t n);

/*
 * Readout: out[k*nby*nbx + by*nbx + bx] = sum of m[y][x] * w[k][y][x] over the cells of bin
 * (bx, by), i.e. columns x_start[bx] <= x < x_start[bx+1] and rows y_start[by] <= y < y_start[by+1]
 * (device arrays of nbx+1 and nby+1 ascending entries from 0 to nx / ny). w: device [nw][ny][nx]
 * FP64 weights, or NULL with nw = 1 for weight 1. Each sum is taken in a fixed order, so it is the
 * same on every call. Replaces the NumPy sums behind Magnets.m_avg_x / m_avg_y / m_avg
 * (core.py:413-424), RegionalOutputReader.read_state (io.py:335-380) and
 * OOPSquareChessOutputReader.read_state (io.py:404-445). scratch: device, of
 * hs_binned_sums_scratch_bytes bytes. Asynchronous.
 */
int hs_binned_sums(const hs_instance_t *inst, int32_t nbx, int32_t nby, const int32_t *x_start,
                   const int32_t *y_start, const double *w, int32_t nw, double *scratch, int64_t scratch_bytes,
                   double *out, void *stream);
int64_t hs_binned_sums_scratch_bytes(const hs_instance_t *inst, int32_t nbx, int32_t nw);

/*
 * Per-energy FP64 fields for the host API (Energy.E / Energy.E_perp): out[0..3] planes =
 * Zeeman E, Zeeman E_perp, exchange E, exchange E_perp (closed forms, energies.py:106-118,
 * 771-781) and out[4], out[5] = combined E and E_perp (value - compensation). Dipolar fields
 * follow as combined minus Zeeman minus exchange. out: [6][ny][nx] doubles.
 */
int hs_field_components(const hs_instance_t *inst, double *out, void *stream);

/*
 * Total energy sum over energies of Energy.E_tot (energies.py:128-130, 368-370, 789-791;
 * core.py:426-428) from the incremental fields, FP64: out[0] = E_tot. Blocking.
 */
int hs_energy_total(const hs_instance_t *inst, double *out_host, void *stream);

/*
 * Swap the Zeeman contribution of the combined fields from the old uniform field (Bx0,By0)
 * to the instance's current field (inst->zee_Bx/By or zee_bx/zee_by arrays):
 * E += m*(E_factor_new - E_factor_old). Replaces ZeemanEnergy.set_field (energies.py:101-114).
 * bx_old/by_old arrays (NULL = uniform old values).
 */
int hs_zeeman_swap(const hs_instance_t *inst, double Bx0, double By0, const float *bx_old,
                   const float *by_old, void *stream);

/*
 * Neel-Arrhenius steps (Magnets._update_Neel core.py:679-709 with E_barrier core.py:625-656
 * and DipolarEnergy/ZeemanEnergy/ExchangeEnergy.update_single energies.py:120-122, 310-329,
 * 783-784), looped as Magnets.progress (core.py:736-760): up to run->max_steps steps while
 * lower_than(t-t0, t_max) and lower_than(MCsteps-MCsteps0, MCsteps_max).
 * state: device hs_state_t[2] (copy 0 holds the input state; the result is in copy
 * (steps_launched+1)&1, reported through *final_copy). partials: device scratch of
 * 32*nblocks*2 bytes, plus 4*nblocks when uniform_beta = 0. log_idx/log_dt: optional device
 * arrays [max_steps] (flat winner index or -1 when no switch; dt added to t). Needs beta64 > 0
 * with uniform_beta = 1; with a temperature that varies over the grid (uniform_beta = 0) it
 * needs beta64c and takes two launches per step (the grid's minimum barrier reorders the
 * waiting times, core.py:688-699). Non-blocking.
 */
int hs_neel_run(const hs_instance_t *inst, const hs_run_t *run, hs_state_t *state, void *partials,
                int32_t *log_idx, double *log_dt, int32_t *final_copy, void *stream);

/*
 * Metropolis steps with multisampling (Magnets._update_Metropolis core.py:659-677, select /
 * _select_grid / _select_single core.py:456-558, update_single / update_multiple
 * energies.py:310-366 with its convolution-or-sum branch rule, ZeemanEnergy / ExchangeEnergy
 * updates), looped as Magnets.progress. ws: device scratch of hs_metropolis_ws_bytes(inst, sel)
 * bytes. state: device hs_state_t[2] (copy 0 is the state). log_cells/log_count: optional
 * per-step accepted cells (capacity sel->nsc per step, at least 1) and counts. A call launches
 * run->max_steps steps and completes at most that many (state.steps_done): the stop predicate
 * can end it earlier, and a presampled step without PBC or log whose selection needs another
 * attempt (core.py:487) takes the next launch for it. Non-blocking.
 */
int hs_metropolis_run(const hs_instance_t *inst, const hs_select_t *sel, const hs_run_t *run,
                      hs_state_t *state, void *ws, int32_t *log_cells, int32_t *log_count,
                      void *stream);

/* Scratch bytes hs_metropolis_run needs for this instance and selection geometry. */
int64_t hs_metropolis_ws_bytes(const hs_instance_t *inst, const hs_select_t *sel);

/* Size of the static scratch of hs_metropolis_run_static (per device). */
#define HS_METRO_STATIC_SCRATCH_BYTES (4 << 20)

/*
 * hs_metropolis_run without a workspace argument and without any allocation: the scratch comes
 * from a file-scope device array of HS_METRO_STATIC_SCRATCH_BYTES bytes that each device holds
 * until the process ends (its address is looked up once per device, on the first call there).
 * Same arguments otherwise and bit-identical results. Calls whose hs_metropolis_ws_bytes exceeds
 * that size return HS_ERR_ARG. Every call on a device shares that device's scratch and clears
 * what it uses first, so calls that overlap on one device (on concurrent streams, or from
 * several threads) race; serialise them. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Non-blocking.
 */
int hs_metropolis_run_static(const hs_instance_t *inst, const hs_select_t *sel, const hs_run_t *run,
                             hs_state_t *state, int32_t *log_cells, int32_t *log_count, void *stream);

/*
 * One zero-temperature sweep: Magnets._minimize_all (core.py:785-806), the step of
 * Magnets.relax (core.py:808-845). For each subset index i of order[0..n_order) (host array, the
 * reference's random permutation), the subset is the cells (y0 + k*step_y, x0 + l*step_x) with
 * (y0, x0) = divmod(i, step_x), skipped when cell (y0, x0) is empty. Every magnet of the subset
 * with switch_energy < 0 (with_barrier: switch_energy/2 + E_B < 0) switches, all at once, and
 * the fields are updated as update_multiple (energies.py:331-366, cutoff / reduced_k as
 * SIMULTANEOUS_SWITCHES_CONVOLUTION_OR_SUM_CUTOFF / REDUCED_KERNEL_SIZE). state[0].switches and
 * .attempts count the switches. ws: device scratch of hs_relax_ws_bytes bytes. Asynchronous.
 */
int hs_relax_sweep(const hs_instance_t *inst, const int32_t *order, int32_t n_order, int32_t step_x,
                   int32_t step_y, int32_t with_barrier, int32_t cutoff, int32_t reduced_k, hs_state_t *state,
                   void *ws, int64_t ws_bytes, void *stream);
int64_t hs_relax_ws_bytes(const hs_instance_t *inst, int32_t step_x, int32_t step_y);

/*
 * The scan of one Magnets.minimize iteration (core.py:763-783): the barrier of every cell
 * (switch_energy = -2E, 0 where empty; with_barrier: the smaller E_barrier of each occupied
 * magnet, core.py:625-656), then cells[0..n) = the cells with barrier < -kT in row-major order
 * (numpy.where order) and result = {n, first cell of the smallest barrier (numpy.argmin), that
 * barrier as the bits of an FP64}. cells: device int32[nx*ny]; result: device int64[3].
 * Asynchronous.
 */
int hs_minimize_scan(const hs_instance_t *inst, int32_t with_barrier, int32_t *cells, int64_t *result, void *stream);

/*
 * Switch the k distinct cells cells[0..k) (device array) together and update the fields as
 * Magnets.update_energy(index) does after such a switch (core.py:241-253): update_single for
 * one cell, update_multiple (cutoff / reduced_k as in hs_relax_sweep) for several.
 * state[0].switches and .attempts grow by k. A cell index outside the grid switches nothing and
 * sets state[0].err. ws: device scratch of hs_flip_ws_bytes bytes. Asynchronous.
 */
int hs_flip_cells(const hs_instance_t *inst, const int32_t *cells, int32_t k, int32_t cutoff, int32_t reduced_k,
                  hs_state_t *state, void *ws, int64_t ws_bytes, void *stream);
int64_t hs_flip_ws_bytes(const hs_instance_t *inst);

/*
 * Wolff cluster steps (Magnets._update_Wolff core.py:711-718 with _select_cluster_exchange
 * core.py:588-602), looped as Magnets.progress until run's stop predicate or max_steps: a seed
 * drawn as _select_single draws it (Philox tag 2, index 2), then every iteration each unchecked
 * cell aligned with the seed and next to the cluster joins with probability
 * 1 - exp(-2 J/kT * sum of the NN weights of its cluster neighbours), on the Philox uniform
 * (cell, step, tag 7 | iteration << 8); the whole cluster switches and the fields follow as
 * update_single / update_multiple (cutoff / reduced_k as in hs_relax_sweep). Needs an exchange
 * energy. run->log: log_cells[0..k) = the cluster (row-major), log_count[0] = k (one step only).
 * ws: device scratch of hs_wolff_ws_bytes bytes. Asynchronous.
 */
int hs_wolff_run(const hs_instance_t *inst, const hs_run_t *run, int32_t cutoff, int32_t reduced_k, hs_state_t *state,
                 void *ws, int64_t ws_bytes, int32_t *log_cells, int32_t *log_count, void *stream);
int64_t hs_wolff_ws_bytes(const hs_instance_t *inst);

/*
 * Neel-Arrhenius steps for many instances in ONE persistent cooperative launch (same
 * semantics per instance as hs_neel_run: core.py:679-709 looped as core.py:736-760, and
 * bit-identical results). Each team keeps its magnets' fields in registers and draws the same
 * Philox streams; a team of several workgroups exchanges one tagged record per workgroup per
 * step (its best candidate and that cell's waiting time) and polls for the round's records
 * (bounded spin, err = 3 on timeout); teams never wait for each other. ws: device scratch of
 * hs_neel_teams_ws_bytes bytes. Asynchronous (descriptors copied through pinned staging; the
 * team array may be released when the call returns). Grids need ny * 2^ceil(log2 nx) <= 2^24.
 * Every instance needs uniform_beta = 1 and beta64 > 0; teams of one call agree on use_perp.
 */
int hs_neel_run_teams(const hs_team_t *teams, int32_t nteams, void *ws, int64_t ws_bytes, void *stream);
int64_t hs_neel_teams_ws_bytes(const hs_team_t *teams, int32_t nteams);

/*
 * Metropolis steps for many instances at once (same semantics per instance as
 * hs_metropolis_run: core.py:659-677 with select core.py:456-558 and the update_multiple
 * branch rule energies.py:331-366, looped as core.py:736-760, and bit-identical results):
 * every step is one selection launch with a workgroup per instance and one update launch over
 * the output tiles of all instances; each instance stops on its own predicate. team.sel is
 * required; team.log_cells / team.log_idx are the optional accepted-cell / count logs of
 * hs_metropolis_run; nblocks and log_dt are unused. ws: device scratch of
 * hs_metro_teams_ws_bytes bytes. Asynchronous (descriptors copied through pinned staging; the
 * team array may be released when the call returns).
 */
int hs_metro_run_teams(const hs_team_t *teams, int32_t nteams, void *ws, int64_t ws_bytes, void *stream);
int64_t hs_metro_teams_ws_bytes(const hs_team_t *teams, int32_t nteams);

/*
 * Latency floor of the persistent Neel step: G co-resident workgroups (one wave each) exchange a
 * tagged candidate per round through the records, stores and bounded polls of k_neel_team and
 * reduce them, `rounds` times in one cooperative launch on the null stream; *us_per_round is the
 * launch's duration (HIP events) over rounds, *failed_rounds (optional) the rounds whose records
 * did not all arrive within the kernel's bounded spin. Allocates and frees its own records and
 * synchronises the device: a measurement, not a stepping call. No counterpart in the reference.
 */
int hs_exchange_round_us(int32_t G, int32_t rounds, double *us_per_round, int32_t *failed_rounds);

/* Steps per captured hipGraph in the step loop of hs_metro_run_teams by default: plain launches
 * (the loop is bound by the GPU, not by the host's launch rate, and replaying graphs measured no
 * faster on MI355X). */
#define HS_GRAPH_STEPS_DEFAULT 0

/*
 * Sets how many steps of the hs_metro_run_teams step loop are captured into one hipGraph and
 * replayed as a unit (rounded down to an even number; 0 or 1: plain launches, one per kernel).
 * A graph is captured the first time a workspace layout and step count are seen and replayed on
 * later calls; the step kernels, their order and their arguments are the same either way, so
 * results are bit-identical. When the caller's stream is itself being captured the loop always
 * launches plainly (the caller's capture records it). Returns the previous setting. No
 * counterpart in the reference (launch mechanics of the device loop).
 */
int hs_set_graph_steps(int32_t steps);

/*
 * Destroys every cached step graph and the private capture streams (after draining the devices
 * they ran on); later graph-mode calls capture again. Measured on MI355X (ROCm 7.2): while
 * executable graphs are alive, plain kernel launches of the process run slower, so a caller that
 * replays graphs in one phase and launches plainly in the next clears the cache in between.
 * Blocking. No counterpart in the reference (launch mechanics of the device loop).
 */
int hs_graph_cache_clear(void);

#ifdef __cplusplus
}
#endif

#endif /* HOTSPICE_HIP_H */
