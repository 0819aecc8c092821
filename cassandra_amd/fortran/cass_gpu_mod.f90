!------------------------------------------------------------------------------
! cass_gpu_mod -- Fortran 2003 iso_c_binding interface to libcassgpu.so
! (include/cass_gpu.h).  A Cassandra build links this module and calls the
! wrappers from the subroutines they replace (energy_routines.f90,
! fragment_growth.f90, widom_insert.f90).  Indices are 1-based here (as in Cassandra) and converted to the
! C ABI's 0-based ones; LOGICAL results come back as default LOGICAL.
!------------------------------------------------------------------------------
MODULE cass_gpu_mod
  USE ISO_C_BINDING
  IMPLICIT NONE

  INTEGER(C_INT32_T), PARAMETER :: CASS_OK = 0
  INTEGER(C_INT32_T), PARAMETER :: CASS_PREC_FP64 = 0, CASS_PREC_MIXED = 1, CASS_PREC_FP32 = 2
  INTEGER(C_INT32_T), PARAMETER :: CASS_MOVE_TRANSLATION = 0, CASS_MOVE_ROTATION = 1, &
       CASS_MOVE_INTRA = 2, CASS_MOVE_INSERTION = 3, CASS_MOVE_DELETION = 4

  ! cass_params (layout of the C struct)
  TYPE, BIND(C) :: cass_params_t
     INTEGER(C_INT32_T) :: box_shape, vdw_style, vdw_sum_style, charge_style
     REAL(C_DOUBLE) :: cell(9)
     REAL(C_DOUBLE) :: rcut_vdw, rcut_coul, rcut_cbmc, rcut_low
     REAL(C_DOUBLE) :: ron_switch, roff_switch, ewald_tol, alpha_dsf_in, temperature
     REAL(C_DOUBLE) :: length_inv(9), hlength(3), basis_length(3), volume
     REAL(C_DOUBLE) :: rcut_vdwsq, rcut_coulsq, rcut_cbmcsq, rcut_lowsq, rcut_max
     REAL(C_DOUBLE) :: ron_switch_sq, roff_switch_sq, switch_factor1, switch_factor2
     REAL(C_DOUBLE) :: rcut3, rcut9, alpha_ewald, h_ewald_cut
     REAL(C_DOUBLE) :: alpha_dsf, dsf_factor1, dsf_factor2, beta
  END TYPE cass_params_t

  TYPE, BIND(C) :: cass_topology_t
     INTEGER(C_INT32_T) :: ntypes
     TYPE(C_PTR) :: eps, sig, mie_n, mie_m
     INTEGER(C_INT32_T) :: nspecies
     TYPE(C_PTR) :: species_natoms, species_atom_offset, species_pair_offset, species_coul_cbmc
     TYPE(C_PTR) :: atom_type, atom_charge, atom_mass, vdw_intra_scale, charge_intra_scale, bonded
  END TYPE cass_topology_t

  TYPE, BIND(C) :: cass_widom_result_t
     REAL(C_DOUBLE) :: widom_sum
     INTEGER(C_INT64_T) :: n_overlaps, n_insertions
     REAL(C_DOUBLE) :: e_const
  END TYPE cass_widom_result_t

  ! rranf state (random_generators.f90): five 64-bit LFSR components
  TYPE, BIND(C) :: cass_rng_t
     INTEGER(C_INT64_T) :: s(5)
  END TYPE cass_rng_t

  TYPE, BIND(C) :: cass_energy_t
     REAL(C_DOUBLE) :: total, intra, inter, bond, angle, dihedral, improper
     REAL(C_DOUBLE) :: intra_vdw, intra_q, inter_vdw, lrc, inter_q, reciprocal, self
  END TYPE cass_energy_t

  INTERFACE
     INTEGER(C_INT32_T) FUNCTION cass_params_derive(p) BIND(C, NAME='cass_params_derive')
       IMPORT :: C_INT32_T, cass_params_t
       TYPE(cass_params_t), INTENT(INOUT) :: p
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_engine_create(device, topo, max_boxes, eng) &
          BIND(C, NAME='cass_engine_create')
       IMPORT :: C_INT32_T, C_PTR, cass_topology_t
       INTEGER(C_INT32_T), VALUE :: device, max_boxes
       TYPE(cass_topology_t), INTENT(IN) :: topo
       TYPE(C_PTR), INTENT(OUT) :: eng
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_engine_destroy(eng) BIND(C, NAME='cass_engine_destroy')
       IMPORT :: C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE :: eng
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_set_precision(eng, prec) BIND(C, NAME='cass_set_precision')
       IMPORT :: C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: prec
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_box_create(eng, p, max_mols, box) BIND(C, NAME='cass_box_create')
       IMPORT :: C_INT32_T, C_PTR, cass_params_t
       TYPE(C_PTR), VALUE :: eng
       TYPE(cass_params_t), INTENT(IN) :: p
       INTEGER(C_INT32_T), VALUE :: max_mols
       INTEGER(C_INT32_T), INTENT(OUT) :: box
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_box_set_params(eng, box, p) BIND(C, NAME='cass_box_set_params')
       IMPORT :: C_INT32_T, C_PTR, cass_params_t
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box
       TYPE(cass_params_t), INTENT(IN) :: p
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_box_upload(eng, box, nmols, mol_species, live, xyz, com, max_dcom) &
          BIND(C, NAME='cass_box_upload')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box, nmols
       INTEGER(C_INT32_T), INTENT(IN) :: mol_species(*), live(*)
       REAL(C_DOUBLE), INTENT(IN) :: xyz(*)
       TYPE(C_PTR), VALUE :: com, max_dcom
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_mol_update(eng, box, mol, live, xyz, com, max_dcom) &
          BIND(C, NAME='cass_mol_update')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box, mol, live
       REAL(C_DOUBLE), INTENT(IN) :: xyz(*)
       TYPE(C_PTR), VALUE :: com, max_dcom
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_mol_inter_energy_c(eng, box, mol, cbmc, xyz, com, max_dcom, &
          e_vdw, e_qq, overlap) BIND(C, NAME='cass_mol_inter_energy')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng, xyz, com, max_dcom
       INTEGER(C_INT32_T), VALUE :: box, mol, cbmc
       REAL(C_DOUBLE), INTENT(OUT) :: e_vdw, e_qq
       INTEGER(C_INT32_T), INTENT(OUT) :: overlap
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_total_energy(eng, box, intra_flag, e, overlap) &
          BIND(C, NAME='cass_total_energy')
       IMPORT :: C_INT32_T, C_PTR, cass_energy_t
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box, intra_flag
       TYPE(cass_energy_t), INTENT(OUT) :: e
       INTEGER(C_INT32_T), INTENT(OUT) :: overlap
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_ewald_setup(eng, box, nvecs) BIND(C, NAME='cass_ewald_setup')
       IMPORT :: C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box
       INTEGER(C_INT32_T), INTENT(OUT) :: nvecs
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_ewald_full(eng, box, e_recip) BIND(C, NAME='cass_ewald_full')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box
       REAL(C_DOUBLE), INTENT(OUT) :: e_recip
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_ewald_move(eng, box, mol, move_flag, xyz, e_recip, de_recip) &
          BIND(C, NAME='cass_ewald_move')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box, mol, move_flag
       REAL(C_DOUBLE), INTENT(IN) :: xyz(*)
       REAL(C_DOUBLE), INTENT(OUT) :: e_recip, de_recip
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_ewald_commit(eng, box) BIND(C, NAME='cass_ewald_commit')
       IMPORT :: C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_ewald_revert(eng, box) BIND(C, NAME='cass_ewald_revert')
       IMPORT :: C_INT32_T, C_PTR
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_trial_energies(eng, box, species, ntrial, trial_xyz, cbmc, exclude, &
          nrg, overlap) BIND(C, NAME='cass_trial_energies')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box, species, ntrial, cbmc, exclude
       REAL(C_DOUBLE), INTENT(IN) :: trial_xyz(*)
       REAL(C_DOUBLE), INTENT(OUT) :: nrg(*)
       INTEGER(C_INT32_T), INTENT(OUT) :: overlap(*)
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_set_energy_table(eng, box, res, rcut_cbmcsq_max) &
          BIND(C, NAME='cass_set_energy_table')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box, res
       REAL(C_DOUBLE), VALUE :: rcut_cbmcsq_max
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_rminsq_table(eng, box, u_max_base, rminsq) &
          BIND(C, NAME='cass_rminsq_table')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box
       REAL(C_DOUBLE), VALUE :: u_max_base
       REAL(C_DOUBLE), INTENT(OUT) :: rminsq(*)
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_widom_set_rminsq(eng, box, rminsq) &
          BIND(C, NAME='cass_widom_set_rminsq')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box
       REAL(C_DOUBLE), INTENT(IN) :: rminsq(*)
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_widom_set_template(eng, box, species, xyz, e_intra_const) &
          BIND(C, NAME='cass_widom_set_template')
       IMPORT :: C_INT32_T, C_PTR, C_DOUBLE
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box, species
       REAL(C_DOUBLE), INTENT(IN) :: xyz(*)
       REAL(C_DOUBLE), INTENT(OUT) :: e_intra_const
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_widom_batch(eng, box, species, n_ins, seed, widom_var, res) &
          BIND(C, NAME='cass_widom_batch')
       IMPORT :: C_INT32_T, C_INT64_T, C_PTR, cass_widom_result_t
       TYPE(C_PTR), VALUE :: eng
       INTEGER(C_INT32_T), VALUE :: box, species
       INTEGER(C_INT64_T), VALUE :: n_ins, seed
       TYPE(C_PTR), VALUE :: widom_var
       TYPE(cass_widom_result_t), INTENT(OUT) :: res
     END FUNCTION
     SUBROUTINE cass_rng_init(r, i1, i3) BIND(C, NAME='cass_rng_init')
       IMPORT :: C_INT64_T, cass_rng_t
       TYPE(cass_rng_t), INTENT(OUT) :: r
       INTEGER(C_INT64_T), VALUE :: i1, i3
     END SUBROUTINE
     REAL(C_DOUBLE) FUNCTION cass_rng_uniform(r) BIND(C, NAME='cass_rng_uniform')
       IMPORT :: C_DOUBLE, cass_rng_t
       TYPE(cass_rng_t), INTENT(INOUT) :: r
     END FUNCTION
     INTEGER(C_INT32_T) FUNCTION cass_accept_or_reject(r, ln_pacc) BIND(C, NAME='cass_accept_or_reject')
       IMPORT :: C_INT32_T, C_DOUBLE, cass_rng_t
       TYPE(cass_rng_t), INTENT(INOUT) :: r
       REAL(C_DOUBLE), VALUE :: ln_pacc
     END FUNCTION
  END INTERFACE

CONTAINS

  ! Drop-in for Compute_Molecule_Nonbond_Inter_Energy(im, is, E_vdw, E_qq,
  ! overlap) (energy_routines.f90:1088) once the molecule's coordinates are
  ! mirrored: lm is the 1-based molecule slot of the engine's box.
  SUBROUTINE Cass_Molecule_Nonbond_Inter_Energy(eng, ibox, lm, cbmc, e_vdw, e_qq, overlap, status)
    TYPE(C_PTR), INTENT(IN) :: eng
    INTEGER, INTENT(IN) :: ibox, lm
    LOGICAL, INTENT(IN) :: cbmc
    REAL(C_DOUBLE), INTENT(OUT) :: e_vdw, e_qq
    LOGICAL, INTENT(OUT) :: overlap
    INTEGER, INTENT(OUT) :: status
    INTEGER(C_INT32_T) :: ov, cb
    cb = 0
    IF (cbmc) cb = 1
    status = cass_mol_inter_energy_c(eng, INT(ibox - 1, C_INT32_T), INT(lm - 1, C_INT32_T), cb, &
         C_NULL_PTR, C_NULL_PTR, C_NULL_PTR, e_vdw, e_qq, ov)
    overlap = ov /= 0
  END SUBROUTINE Cass_Molecule_Nonbond_Inter_Energy

END MODULE cass_gpu_mod
