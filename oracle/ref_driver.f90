! ref_driver.f90 -- C-callable driver for the reference pysic Fortran core.
!
! TEST INFRASTRUCTURE ONLY.  build_ref.sh compiles the reference's own
! Fortran sources where they lie (they are never copied here) together with
! this driver into oracle/_ref/libpysic_ref.so.  The tests use it to write
! golden vectors and to pin the C restatement in pysic_oracle.c; bench.py
! may time it as the CPU baseline.  Nothing in the product links it.
!
! Each entry point replays one step of the Python-side set-up order of the
! reference (pysic/calculator.py:1595-1637 initialize_fortran_core and
! :1063-1446 update_core_potentials) through the f2py interface module
! (fortran/PyInterface.f90).  Element labels use the reference's two-integer
! code (pysic/utility/f2py.py: 'S'=119, 'i'=9, ' '=0).
module pysic_ref_capi
  use iso_c_binding
  use pysic_interface
  implicit none
contains

  subroutine ref_start() bind(C, name="ref_start")
    ! pysic/core.py:20-28
    call start_potentials()
    call start_bond_order_factors()
    call start_mpi()
    call start_rng(5301)
  end subroutine ref_start

  subroutine ref_create_system(n, masses, charges, positions, momenta, tags, elements, &
       vectors, inverse, pbc) bind(C, name="ref_create_system")
    ! PyInterface.f90:250 create_atoms, :116 distribute_mpi, :208 create_cell
    integer(c_int), value :: n
    real(c_double), intent(in) :: masses(n), charges(n), positions(3,n), momenta(3,n)
    integer(c_int), intent(in) :: tags(n), elements(2,n), pbc(3)
    real(c_double), intent(in) :: vectors(3,3), inverse(3,3)
    logical :: periodic(3)
    integer :: i
    call create_atoms(n, masses, charges, positions, momenta, tags, elements)
    call distribute_mpi(n)
    do i = 1, 3
       periodic(i) = pbc(i) /= 0
    end do
    call create_cell(vectors, inverse, periodic)
  end subroutine ref_create_system

  subroutine ref_update_coordinates(n, positions, momenta) bind(C, name="ref_update_coordinates")
    integer(c_int), value :: n
    real(c_double), intent(in) :: positions(3,n), momenta(3,n)
    call update_atom_coordinates(n, positions, momenta)
  end subroutine ref_update_coordinates

  subroutine ref_update_charges(n, charges) bind(C, name="ref_update_charges")
    integer(c_int), value :: n
    real(c_double), intent(in) :: charges(n)
    call update_atom_charges(n, charges)
  end subroutine ref_update_charges

  subroutine ref_allocate_potentials(npot) bind(C, name="ref_allocate_potentials")
    integer(c_int), value :: npot
    call allocate_potentials(npot)
  end subroutine ref_allocate_potentials

  ! kind: the potential type codes of the engine's descriptors
  ! (1 exponential, 2 charge_self, 3 LJ, 4 spring, 5 Buckingham, 6 power,
  ! 7 tabulated, 8 shift_power, 9 step, 10 constant, 11 force).
  ! is_mult /= 0 adds a ProductPotential multiplier, kept until
  ! ref_clear_multipliers (calculator.py:1262-1277, 1368-1369).
  function ref_add_potential(kind, n_targets, n_params, params, cutoff, soft, elements, &
       group, is_mult) result(ok) bind(C, name="ref_add_potential")
    integer(c_int), value :: kind, n_targets, n_params, group, is_mult
    real(c_double), intent(in) :: params(n_params)
    real(c_double), value :: cutoff, soft
    integer(c_int), intent(in) :: elements(2, n_targets)
    integer(c_int) :: ok
    integer :: notags(n_targets), noinds(n_targets)
    logical :: success, mult
    notags = -9
    noinds = -9
    mult = is_mult /= 0
    success = .false.
    if (kind >= 1 .and. kind <= 15) then
       call add_potential(n_targets, n_params, trim(pot_name(kind)), params, cutoff, soft, &
            elements, notags, noinds, elements, notags, noinds, group, mult, success)
    end if
    ok = 0
    if (success) ok = 1
  end function ref_add_potential

  ! The same with the original targets given apart from the applied ones (a
  ! permuted copy keeps its potential's original elements, calculator.py:
  ! 1241-1277); the charge-dependent pair terms read them.
  function ref_add_potential_orig(kind, n_targets, n_params, params, cutoff, soft, elements, &
       orig_elements, group, is_mult) result(ok) bind(C, name="ref_add_potential_orig")
    integer(c_int), value :: kind, n_targets, n_params, group, is_mult
    real(c_double), intent(in) :: params(n_params)
    real(c_double), value :: cutoff, soft
    integer(c_int), intent(in) :: elements(2, n_targets), orig_elements(2, n_targets)
    integer(c_int) :: ok
    integer :: notags(n_targets), noinds(n_targets)
    logical :: success, mult
    notags = -9
    noinds = -9
    mult = is_mult /= 0
    success = .false.
    if (kind >= 1 .and. kind <= 15) then
       call add_potential(n_targets, n_params, trim(pot_name(kind)), params, cutoff, soft, &
            elements, notags, noinds, orig_elements, notags, noinds, group, mult, success)
    end if
    ok = 0
    if (success) ok = 1
  end function ref_add_potential_orig

  function pot_name(kind) result(name)
    integer(c_int), intent(in) :: kind
    character(len=11) :: name
    character(len=11) :: names(15)
    names = (/ 'exponential', 'charge_self', 'LJ         ', 'spring     ', 'Buckingham ', &
         'power      ', 'tabulated  ', 'shift_power', 'step       ', 'constant   ', &
         'force      ', 'charge_exp ', 'charge_abs ', 'charge_pair', 'bond_bend  ' /)
    name = names(kind)
  end function pot_name

  subroutine ref_clear_multipliers() bind(C, name="ref_clear_multipliers")
    call clear_potential_multipliers()
  end subroutine ref_clear_multipliers

  subroutine ref_allocate_bond_order_factors(nb) bind(C, name="ref_allocate_bond_order_factors")
    integer(c_int), value :: nb
    call allocate_bond_order_factors(nb)
  end subroutine ref_allocate_bond_order_factors

  ! 'tersoff' with params [beta, eta, a, c, d, h, mu] split (2, 5)
  function ref_add_tersoff(params, cutoff, soft, elements, group) result(ok) &
       bind(C, name="ref_add_tersoff")
    real(c_double), intent(in) :: params(7)
    real(c_double), value :: cutoff, soft
    integer(c_int), intent(in) :: elements(2,3)
    integer(c_int), value :: group
    integer(c_int) :: ok
    integer :: split(2)
    logical :: success
    split = (/ 2, 5 /)
    call add_bond_order_factor(3, 7, 2, 'tersoff', params, split, cutoff, soft, &
         elements, elements, group, success)
    ok = 0
    if (success) ok = 1
  end function ref_add_tersoff

  subroutine ref_allocate_bond_order_storage(n, ngroups, nfactors) &
       bind(C, name="ref_allocate_bond_order_storage")
    integer(c_int), value :: n, ngroups, nfactors
    call allocate_bond_order_storage(n, ngroups, nfactors)
  end subroutine ref_allocate_bond_order_storage

  subroutine ref_generate_neighbor_lists(n, cutoffs) bind(C, name="ref_generate_neighbor_lists")
    integer(c_int), value :: n
    real(c_double), intent(in) :: cutoffs(n)
    call generate_neighbor_lists(n, cutoffs)
  end subroutine ref_generate_neighbor_lists

  subroutine ref_neighbor_count(i, nn) bind(C, name="ref_neighbor_count")
    integer(c_int), value :: i
    integer(c_int), intent(out) :: nn
    integer :: tmp
    call get_number_of_neighbors_of_atom(i, tmp)
    nn = tmp
  end subroutine ref_neighbor_count

  subroutine ref_neighbor_list(i, nn, nbrs, offsets) bind(C, name="ref_neighbor_list")
    integer(c_int), value :: i, nn
    integer(c_int), intent(out) :: nbrs(nn), offsets(3,nn)
    call get_neighbor_list_of_atom(i, nn, nbrs, offsets)
  end subroutine ref_neighbor_list

  subroutine ref_create_lists() bind(C, name="ref_create_lists")
    call create_potential_list()
    call create_bond_order_factor_list()
  end subroutine ref_create_lists

  subroutine ref_set_ewald(n, rc, kc, kcells, sigma, eps, scaler) bind(C, name="ref_set_ewald")
    integer(c_int), value :: n
    real(c_double), value :: rc, kc, sigma, eps
    integer(c_int), intent(in) :: kcells(3)
    real(c_double), intent(in) :: scaler(n)
    integer :: kk(3)
    kk = kcells
    call set_ewald_parameters(n, rc, kc, kk, sigma, eps, scaler)
  end subroutine ref_set_ewald

  subroutine ref_energy(e) bind(C, name="ref_energy")
    real(c_double), intent(out) :: e
    call calculate_energy(e)
  end subroutine ref_energy

  subroutine ref_forces(n, forces, stress) bind(C, name="ref_forces")
    integer(c_int), value :: n
    real(c_double), intent(out) :: forces(3,n), stress(6)
    call calculate_forces(n, forces, stress)
  end subroutine ref_forces

  subroutine ref_electronegativities(n, enegs) bind(C, name="ref_electronegativities")
    integer(c_int), value :: n
    real(c_double), intent(out) :: enegs(n)
    call calculate_electronegativities(n, enegs)
  end subroutine ref_electronegativities

  ! --- descriptor registry (PyInterface.f90:706-1011) -------------------
  ! Names travel as the reference's integer codes (Utility.f90:205-250).

  subroutine ref_number_of_types(npot, nbond) bind(C, name="ref_number_of_types")
    integer(c_int), intent(out) :: npot, nbond
    integer :: a, b
    call number_of_potentials(a)
    call number_of_bond_order_factors(b)
    npot = a
    nbond = b
  end subroutine ref_number_of_types

  subroutine ref_list_potentials(npot, codes) bind(C, name="ref_list_potentials")
    integer(c_int), value :: npot
    integer(c_int), intent(out) :: codes(11, npot)
    integer :: tmp(11, npot)
    call list_valid_potentials(npot, tmp)
    codes = tmp
  end subroutine ref_list_potentials

  subroutine ref_list_bond_order_factors(nbond, codes) bind(C, name="ref_list_bond_order_factors")
    integer(c_int), value :: nbond
    integer(c_int), intent(out) :: codes(11, nbond)
    integer :: tmp(11, nbond)
    call list_valid_bond_order_factors(nbond, tmp)
    codes = tmp
  end subroutine ref_list_bond_order_factors

  function c_name(n, chars) result(name)
    integer(c_int), intent(in) :: n
    character(kind=c_char), intent(in) :: chars(n)
    character(len=n) :: name
    integer :: i
    do i = 1, n
       name(i:i) = chars(i)
    end do
  end function c_name

  ! targets, parameter count and parameter-name codes of a potential
  subroutine ref_potential_info(n, chars, n_targets, n_params, names) &
       bind(C, name="ref_potential_info")
    integer(c_int), value :: n
    character(kind=c_char), intent(in) :: chars(n)
    integer(c_int), intent(out) :: n_targets, n_params, names(10, 12)
    integer :: a, b, tmp(10, 12)
    call number_of_targets_of_potential(c_name(n, chars), a)
    call number_of_parameters_of_potential(c_name(n, chars), b)
    call names_of_parameters_of_potential(c_name(n, chars), tmp)
    n_targets = a
    n_params = b
    names = tmp
  end subroutine ref_potential_info

  ! targets, level, parameter counts and names per target class (1, 2)
  subroutine ref_bond_order_factor_info(n, chars, n_targets, level, n_params, names) &
       bind(C, name="ref_bond_order_factor_info")
    integer(c_int), value :: n
    character(kind=c_char), intent(in) :: chars(n)
    integer(c_int), intent(out) :: n_targets, level, n_params(2), names(10, 12, 2)
    integer :: a, b, c, k, tmp(10, 12)
    call number_of_targets_of_bond_order_factor(c_name(n, chars), a)
    call level_of_bond_order_factor(c_name(n, chars), b)
    n_targets = a
    level = b
    do k = 1, 2
       call number_of_parameters_of_bond_order_factor(c_name(n, chars), k, c)
       n_params(k) = c
       call names_of_parameters_of_bond_order_factor(c_name(n, chars), k, tmp)
       names(:, :, k) = tmp
    end do
  end subroutine ref_bond_order_factor_info

  ! --- bond-order factor queries (PyInterface.f90:579-634) ---------------
  subroutine ref_bond_order_factors(n, group, factors) bind(C, name="ref_bond_order_factors")
    integer(c_int), value :: n, group
    real(c_double), intent(out) :: factors(n)
    call calculate_bond_order_factors(n, group, factors)
  end subroutine ref_bond_order_factors

  ! atom: 1-based, as bondorder.py:518 passes it
  subroutine ref_bond_order_gradients(n, group, atom, grads) bind(C, name="ref_bond_order_gradients")
    integer(c_int), value :: n, group, atom
    real(c_double), intent(out) :: grads(3, n)
    call calculate_bond_order_gradients(n, group, atom, grads)
  end subroutine ref_bond_order_gradients

  subroutine ref_bond_order_gradients_of_factor(n, group, atom, grads) &
       bind(C, name="ref_bond_order_gradients_of_factor")
    integer(c_int), value :: n, group, atom
    real(c_double), intent(out) :: grads(3, n)
    call calculate_bond_order_gradients_of_factor(n, group, atom, grads)
  end subroutine ref_bond_order_gradients_of_factor

  ! --- the stand-alone Ewald energy (PyInterface.f90:1050) ----------------
  subroutine ref_get_ewald_energy(rc, kc, kcells, sigma, eps, e) bind(C, name="ref_get_ewald_energy")
    real(c_double), value :: rc, kc, sigma, eps
    integer(c_int), intent(in) :: kcells(3)
    real(c_double), intent(out) :: e
    integer :: kk(3)
    kk = kcells
    call get_ewald_energy(rc, kc, kk, sigma, eps, e)
  end subroutine ref_get_ewald_energy

  subroutine ref_release() bind(C, name="ref_release")
    call release()
  end subroutine ref_release

end module pysic_ref_capi
