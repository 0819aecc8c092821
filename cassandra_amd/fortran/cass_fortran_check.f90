!------------------------------------------------------------------------------
! cass_fortran_check -- a Fortran host driving libcassgpu through
! cass_gpu_mod: 125 LJ sites on a jittered lattice in a cubic box (LJ
! cut_tail 8.5 A, 120 K / 3.4 A), total energy and one molecule energy.
! Prints "total <E>" and "mol1 <E_vdw>" for the test to compare with the
! oracle, and the first draws of the rranf stream ("rng1".."rng5").
!------------------------------------------------------------------------------
PROGRAM cass_fortran_check
  USE ISO_C_BINDING
  USE cass_gpu_mod
  IMPLICIT NONE
  TYPE(cass_params_t) :: p
  TYPE(cass_topology_t) :: t
  TYPE(cass_energy_t) :: e
  TYPE(C_PTR) :: eng
  TYPE(cass_rng_t) :: rng
  INTEGER(C_INT32_T) :: rc, box, ov
  INTEGER, PARAMETER :: n = 125
  REAL(C_DOUBLE), TARGET :: eps(1), sig(1), charge(1), mass(1), vsc(1), qsc(1)
  INTEGER(C_INT32_T), TARGET :: sp_nat(1), sp_off(1), sp_poff(1), sp_cbmc(1), atype(1), bonded(1)
  INTEGER(C_INT32_T) :: mol_species(n), live(n)
  REAL(C_DOUBLE) :: xyz(3 * n), L, evdw, eqq
  INTEGER :: i, j, k, m
  LOGICAL :: ovl
  INTEGER :: status

  L = 5.0_C_DOUBLE * 3.4_C_DOUBLE * 1.12_C_DOUBLE
  p%box_shape = 0
  p%vdw_style = 1
  p%vdw_sum_style = 1
  p%charge_style = 0
  p%cell = 0.0_C_DOUBLE
  p%cell(1) = L
  p%cell(5) = L
  p%cell(9) = L
  p%rcut_vdw = 8.5_C_DOUBLE
  p%rcut_cbmc = 6.5_C_DOUBLE
  p%rcut_low = 1.0_C_DOUBLE
  p%temperature = 120.0_C_DOUBLE
  rc = cass_params_derive(p)

  eps(1) = 120.0_C_DOUBLE * 0.8314472_C_DOUBLE
  sig(1) = 3.4_C_DOUBLE
  charge(1) = 0.0_C_DOUBLE
  mass(1) = 39.948_C_DOUBLE
  vsc(1) = 0.0_C_DOUBLE
  qsc(1) = 0.0_C_DOUBLE
  sp_nat(1) = 1
  sp_off(1) = 0
  sp_poff(1) = 0
  sp_cbmc(1) = 1
  atype(1) = 0
  bonded(1) = 0
  t%ntypes = 1
  t%eps = C_LOC(eps)
  t%sig = C_LOC(sig)
  t%mie_n = C_NULL_PTR
  t%mie_m = C_NULL_PTR
  t%nspecies = 1
  t%species_natoms = C_LOC(sp_nat)
  t%species_atom_offset = C_LOC(sp_off)
  t%species_pair_offset = C_LOC(sp_poff)
  t%species_coul_cbmc = C_LOC(sp_cbmc)
  t%atom_type = C_LOC(atype)
  t%atom_charge = C_LOC(charge)
  t%atom_mass = C_LOC(mass)
  t%vdw_intra_scale = C_LOC(vsc)
  t%charge_intra_scale = C_LOC(qsc)
  t%bonded = C_LOC(bonded)

  m = 0
  DO i = 0, 4
     DO j = 0, 4
        DO k = 0, 4
           m = m + 1
           xyz(3 * m - 2) = (REAL(i, C_DOUBLE) + 0.5_C_DOUBLE) * L / 5.0_C_DOUBLE - 0.5_C_DOUBLE * L &
                + 0.07_C_DOUBLE * SIN(REAL(m, C_DOUBLE))
           xyz(3 * m - 1) = (REAL(j, C_DOUBLE) + 0.5_C_DOUBLE) * L / 5.0_C_DOUBLE - 0.5_C_DOUBLE * L &
                + 0.07_C_DOUBLE * COS(REAL(3 * m, C_DOUBLE))
           xyz(3 * m) = (REAL(k, C_DOUBLE) + 0.5_C_DOUBLE) * L / 5.0_C_DOUBLE - 0.5_C_DOUBLE * L &
                + 0.07_C_DOUBLE * SIN(REAL(7 * m, C_DOUBLE))
           mol_species(m) = 0
           live(m) = 1
        END DO
     END DO
  END DO

  rc = cass_engine_create(0, t, 1, eng)
  IF (rc /= CASS_OK) STOP 'cass_engine_create failed'
  rc = cass_box_create(eng, p, n, box)
  IF (rc /= CASS_OK) STOP 'cass_box_create failed'
  rc = cass_box_upload(eng, box, n, mol_species, live, xyz, C_NULL_PTR, C_NULL_PTR)
  IF (rc /= CASS_OK) STOP 'cass_box_upload failed'
  rc = cass_total_energy(eng, box, 1, e, ov)
  IF (rc /= CASS_OK .OR. ov /= 0) STOP 'cass_total_energy failed'
  CALL Cass_Molecule_Nonbond_Inter_Energy(eng, 1, 1, .FALSE., evdw, eqq, ovl, status)
  IF (status /= CASS_OK) STOP 'molecule energy failed'
  WRITE(*, '(A,ES26.17E3)') 'total ', e%total
  WRITE(*, '(A,ES26.17E3)') 'mol1 ', evdw
  WRITE(*, '(A,ES26.17E3)') 'box ', L
  ! the reference's rranf stream from the Fortran side (seeds 1234567, 7654321)
  CALL cass_rng_init(rng, 1234567_C_INT64_T, 7654321_C_INT64_T)
  DO i = 1, 5
     WRITE(*, '(A,I0,A,ES26.17E3)') 'rng', i, ' ', cass_rng_uniform(rng)
  END DO
  rc = cass_engine_destroy(eng)
END PROGRAM cass_fortran_check
