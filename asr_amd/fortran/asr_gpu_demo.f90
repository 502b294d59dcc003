!> A short Fortran caller of libasr_hip.so through module asr_gpu, as asr's chain drivers would call it.
!>   asr_gpu_demo                  probe: the device count and a refused call with the engine's message (no GPU work)
!>   asr_gpu_demo <in> <out>       read a site-class operator in the reference's table layout (the fields of augcls_op,
!>                                 1-based, column-major) and a batch of chains, copy the operator with gpu_op_from_tables,
!>                                 run the complex-symmetric edc chains and the real Lanczos chains, write a_n and b_n
!> Input (list-directed): nn nslot nop nsmax ncls nsp ntype nstab; hopr hopg (nslot, nn); flipr flipg (nsmax, nn); lst ncl
!> ntyp onstart (nn); sq (nn); stab_ptr (nn + 1); stab (nstab); pi (9, nop); sg (9, nop); block (9, 9, nslot, ntype);
!> block0 (9, 9, ntype); nsc (ncls); spc (nsmax, ncls); mt (nsmax, nsmax, nsmax, ncls); nchain nlev; start_orb frozen_sp
!> (nchain, 0-based); params (9, nsp, 4, nchain); pab (9, nsp, 2, nchain); z (2, nchain).
program asr_gpu_demo
  use, intrinsic :: iso_c_binding
  use asr_gpu
  implicit none
  character(len=1024) :: fin, fout
  integer :: nn, nslot, nop, nsmax, ncls, nsp, ntype, nstab, nchain, nlev, u, j, k
  integer, allocatable :: hopr(:, :), hopg(:, :), flipr(:, :), flipg(:, :), lst(:), ncl(:), ntyp(:), ist(:), stab_ptr(:), &
                          stab(:), pi(:, :), nsc(:), spc(:, :)
  real(c_double), allocatable :: sq(:), sg(:, :), block(:, :, :, :), block0(:, :, :), mt(:, :, :, :)
  integer(c_int), allocatable :: so(:), fz(:), nl(:), mu(:), mr(:)
  real(c_double), allocatable :: par(:), pab(:), zr(:, :), ar(:, :), br(:, :)
  complex(c_double_complex), allocatable :: z(:), a(:, :), b(:, :)
  real(c_double) :: ms
  type(c_ptr) :: h
  integer(c_int) :: st

  if (command_argument_count() < 2) then
    print '(a, i0)', 'devices ', gpu_device_count()
    st = c_asr_gpu_op_create(10, 18, 65, 3, 1, 3, 1, h)  ! 65 operations: refused before any device work
    print '(a, i0, a, a)', 'status ', st, ' message ', gpu_last_error()
    stop
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=u, file=trim(fin), status='old', action='read')
  read (u, *) nn, nslot, nop, nsmax, ncls, nsp, ntype, nstab
  allocate (hopr(nslot, nn), hopg(nslot, nn), flipr(nsmax, nn), flipg(nsmax, nn), lst(nn), ncl(nn), ntyp(nn), ist(nn), &
            sq(nn), stab_ptr(nn + 1), stab(nstab), pi(9, nop), sg(9, nop), block(9, 9, nslot, ntype), &
            block0(9, 9, ntype), nsc(ncls), spc(nsmax, ncls), mt(nsmax, nsmax, nsmax, ncls))
  read (u, *) hopr, hopg, flipr, flipg, lst, ncl, ntyp, ist, sq, stab_ptr, stab, pi, sg, block, block0, nsc, spc, mt
  read (u, *) nchain, nlev
  allocate (so(nchain), fz(nchain), nl(nchain), mu(nchain), mr(nchain), par(9 * nsp * 4 * nchain), &
            pab(9 * nsp * 2 * nchain), zr(2, nchain), z(nchain), a(nlev, nchain), b(nlev, nchain), ar(nlev, nchain), &
            br(nlev, nchain))
  read (u, *) so, fz, par, pab, zr
  close (u)
  z = cmplx(zr(1, :), zr(2, :), c_double_complex)
  nl = nlev

  h = gpu_op_from_tables(nn, nslot, nop, nsmax, ncls, nsp, ntype, hopr, hopg, flipr, flipg, lst, ncl, ntyp, ist /= 0, sq, &
                         stab_ptr, stab, pi, sg, block, block0, nsc, spc, mt)
  call gpu_check(c_asr_gpu_chains_edc(h, nchain, so, par, pab, fz, c_null_ptr, z, nl, 1, nlev, a, b, mu, 0, ms), &
                 'asr_gpu_chains_edc')
  call gpu_check(c_asr_gpu_lanczos_real(h, nchain, so, par, fz, c_null_ptr, nl, 1.0e-12_c_double, nlev, ar, br, mr, 0, ms), &
                 'asr_gpu_lanczos_real')
  call c_asr_gpu_op_destroy(h)

  open (newunit=u, file=trim(fout), status='replace', action='write')
  do j = 1, nchain
    write (u, '(2i6)') mu(j), mr(j)
    do k = 1, nlev
      write (u, '(6es25.16e3)') real(a(k, j)), aimag(a(k, j)), real(b(k, j)), aimag(b(k, j)), ar(k, j), br(k, j)
    end do
  end do
  close (u)
  print '(a, i0, a, i0, a)', 'chains ', nchain, ' x ', nlev, ' levels written'
end program asr_gpu_demo
