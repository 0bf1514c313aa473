! Caller of the libgint Fortran module (the binding CP2K uses), linked
! against this library: every public entry point of the module, in protocol
! order, for one set quartet of a two-primitive, two-contraction s shell,
! K(4x4) = exchange with P(i) = 0.01 i.  Usage: drive_libgint <table file>.
! Built by the library build (where the reference's module source exists)
! and by tests/test_abi_host.py; run by tests/test_gpu_fortran.py.
program drive_libgint
   use iso_c_binding
   use libgint
   implicit none
   type(c_ptr) :: h
   real(kind=8) :: R(3), Z(2), K(4), cell(3,3), celli(3,3), neighs(3,1), P(16), Kmat(16), C0(105*22, 207)
   character(len=512) :: table
   integer :: i
   R = (/ 0.0d0, 0.0d0, 0.0d0 /)
   Z = (/ 1.3d0, 0.4d0 /)
   K = (/ 0.6d0, 0.5d0, 0.2d0, 0.9d0 /)
   cell = 0.0d0
   celli = 0.0d0
   do i = 1, 3
      cell(i, i) = 100.0d0
      celli(i, i) = 0.01d0
   end do
   neighs = 0.0d0
   Kmat = 0.0d0
   ! the truncated-Coulomb table (207 patches x 22 orders x 105 coefficients)
   call get_command_argument(1, table)
   open(unit=10, file=trim(table), access='stream', form='unformatted', status='old')
   read(10) C0
   close(10)
   do i = 1, 16
      P(i) = 1.0d-2 * i
   end do
   call libgint_init(h)
   call libgint_set_Potential_Truncated(h, 11.3384d0, C0)
   call libgint_set_hf_fac(h, 1.0d0)
   call libgint_set_max_mem(h, 1024)
   call libgint_set_cell(h, .false._c_bool, cell, celli)
   call libgint_set_neighs(h, neighs, 1)
   call libgint_set_Atom(h, 0, R, Z, 2)
   call libgint_set_Atom_L(h, 0, 0, 2, K)
   call libgint_set_P(h, P)
   call libgint_add_prm(h, 0, 0, 0, 0)
   call libgint_add_prm(h, 1, 1, 1, 1)
   call libgint_add_shell(h, 0, 0, 0, 0, 0, 0)
   call libgint_add_cell(h)
   call libgint_add_qrt(h, 0, 0, 0, 0, 2, 2, 2, 2)
   do i = 0, 15
      call libgint_add_qrtt(h, 0.25d0, 0, 0, 0, 0, i / 8, mod(i / 4, 2), mod(i / 2, 2), mod(i, 2), &
                            4, 4, 4, 4, i / 8 + 4 * mod(i / 2, 2), i / 8 + 4 * mod(i, 2), &
                            mod(i / 4, 2) + 4 * mod(i / 2, 2), mod(i / 4, 2) + 4 * mod(i, 2), &
                            .false._c_bool, .false._c_bool, .false._c_bool, .false._c_bool)
   end do
   call libgint_add_set(h)
   call libgint_dispatch(h)
   call libgint_get_K(h, Kmat)
   do i = 1, 16
      print '(es25.16e3)', Kmat(i)
   end do
end program drive_libgint
