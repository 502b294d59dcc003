!> Fortran binding of libasr_hip.so (include/asr_hip.h) for asr's chain drivers: explicit iso_c_binding interfaces of the
!> C ABI, the status check that turns a non-zero status into `error stop` with the engine's message (the reference's
!> error behaviour), and the copy of a site-class operator given in the reference's own table layout (the fields of
!> augcls_op, src/asr_augcls.f90:45-68: 1-based node, operation, orbital and species numbers, 0 = no link, column-major
!> arrays) onto the device.  Modelled on the reference's libxc binding (src/asr_xc.f90:23-55: opaque c_ptr handles, value
!> arguments).  No dependency on the reference's modules: asr_gpu_augcls adds the routines that take augcls_op itself.
module asr_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: gpu_check, gpu_last_error, gpu_op_from_tables, gpu_device_count, c_asr_gpu_op_create
  public :: c_asr_gpu_op_destroy, c_asr_gpu_op_trim, c_asr_gpu_op_set_tile, c_asr_gpu_op_set_depth_tol, &
            c_asr_gpu_op_set_level_form, c_asr_gpu_op_set_window, c_asr_gpu_op_set_window_queue, &
            c_asr_gpu_op_set_compaction, &
            c_asr_gpu_chains_edc, c_asr_gpu_lanczos_real, c_asr_gpu_ed3_blocks, &
            c_asr_gpu_real_create, c_asr_gpu_real_destroy, c_asr_gpu_kpm_moments

  interface
    function c_asr_gpu_last_error() bind(C, name='asr_gpu_last_error') result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function
    function c_asr_gpu_device_count(n) bind(C, name='asr_gpu_device_count') result(ierr)
      import :: c_int
      integer(c_int), intent(out) :: n
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_create(nn, nslot, nop, nsmax, ncls, nsp, ntype, op) bind(C, name='asr_gpu_op_create') result(ierr)
      import :: c_int, c_ptr
      integer(c_int), value :: nn, nslot, nop, nsmax, ncls, nsp, ntype
      type(c_ptr), intent(out) :: op
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_tables(op, hopr, hopg, flipr, flipg, lst, ncl, ntyp, fz, sq, stab_ptr, stab, nstab) &
        bind(C, name='asr_gpu_op_set_tables') result(ierr)
      import :: c_ptr, c_int, c_double, c_signed_char
      type(c_ptr), value :: op
      integer(c_int), intent(in) :: hopr(*), hopg(*), flipr(*), flipg(*), lst(*), ncl(*), ntyp(*), stab_ptr(*), stab(*)
      integer(c_signed_char), intent(in) :: fz(*)
      real(c_double), intent(in) :: sq(*)
      integer(c_int), value :: nstab
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_orbital_ops(op, pi, sg) bind(C, name='asr_gpu_op_set_orbital_ops') result(ierr)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: op
      integer(c_int), intent(in) :: pi(*)
      real(c_double), intent(in) :: sg(*)
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_blocks(op, block, block0) bind(C, name='asr_gpu_op_set_blocks') result(ierr)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: op
      real(c_double), intent(in) :: block(*), block0(*)
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_classes(op, nsc, spc, mt) bind(C, name='asr_gpu_op_set_classes') result(ierr)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: op
      integer(c_int), intent(in) :: nsc(*), spc(*)
      real(c_double), intent(in) :: mt(*)
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_tile(op, mt) bind(C, name='asr_gpu_op_set_tile') result(ierr)
      import :: c_ptr, c_int
      type(c_ptr), value :: op
      integer(c_int), value :: mt
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_depth_tol(op, depth_tol, tail) bind(C, name='asr_gpu_op_set_depth_tol') result(ierr)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: op
      real(c_double), value :: depth_tol
      integer(c_int), value :: tail
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_level_form(op, form) bind(C, name='asr_gpu_op_set_level_form') result(ierr)
      import :: c_ptr, c_int
      type(c_ptr), value :: op
      integer(c_int), value :: form
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_window(op, nslot) bind(C, name='asr_gpu_op_set_window') result(ierr)
      import :: c_ptr, c_int
      type(c_ptr), value :: op
      integer(c_int), value :: nslot
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_window_queue(op, on_device) bind(C, name='asr_gpu_op_set_window_queue') result(ierr)
      import :: c_ptr, c_int
      type(c_ptr), value :: op
      integer(c_int), value :: on_device
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_op_set_compaction(op, on) bind(C, name='asr_gpu_op_set_compaction') result(ierr)
      import :: c_ptr, c_int
      type(c_ptr), value :: op
      integer(c_int), value :: on
      integer(c_int) :: ierr
    end function
    subroutine c_asr_gpu_op_destroy(op) bind(C, name='asr_gpu_op_destroy')
      import :: c_ptr
      type(c_ptr), value :: op
    end subroutine
    function c_asr_gpu_op_trim(op) bind(C, name='asr_gpu_op_trim') result(ierr)
      import :: c_ptr, c_int
      type(c_ptr), value :: op
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_chains_edc(op, nchain, start_orb, params, pab, frozen_sp, conc, z, nlev, project, ldab, a, b, mused, &
                                  mode, elapsed_ms) bind(C, name='asr_gpu_chains_edc') result(ierr)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: op, conc
      integer(c_int), value :: nchain, project, ldab, mode
      integer(c_int), intent(in) :: start_orb(*), frozen_sp(*), nlev(*)
      real(c_double), intent(in) :: params(*), pab(*)
      complex(c_double_complex), intent(in) :: z(*)
      complex(c_double_complex), intent(out) :: a(*), b(*)
      integer(c_int), intent(out) :: mused(*)
      real(c_double), intent(out) :: elapsed_ms
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_lanczos_real(op, nchain, start_orb, params, frozen_sp, conc, nlev, tiny, ldab, a, b, mused, mode, &
                                    elapsed_ms) bind(C, name='asr_gpu_lanczos_real') result(ierr)
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: op, conc
      integer(c_int), value :: nchain, ldab, mode
      integer(c_int), intent(in) :: start_orb(*), frozen_sp(*), nlev(*)
      real(c_double), intent(in) :: params(*)
      real(c_double), value :: tiny
      real(c_double), intent(out) :: a(*), b(*), elapsed_ms
      integer(c_int), intent(out) :: mused(*)
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_ed3_blocks(op, nchain, start_orb, params, cf, oc, frozen_sp, conc, z, nlev, project, ldab, bk, &
                                  a, b, c, mused, mode, elapsed_ms) bind(C, name='asr_gpu_ed3_blocks') result(ierr)
      import :: c_ptr, c_int, c_double, c_double_complex
      type(c_ptr), value :: op, conc
      integer(c_int), value :: nchain, project, ldab, mode
      integer(c_int), intent(in) :: start_orb(*), frozen_sp(*), nlev(*)
      real(c_double), intent(in) :: params(*), cf(*), oc(*)
      complex(c_double_complex), intent(in) :: z(*)
      complex(c_double_complex), intent(out) :: bk(*), a(*), b(*), c(*)
      integer(c_int), intent(out) :: mused(*)
      real(c_double), intent(out) :: elapsed_ms
      integer(c_int) :: ierr
    end function
    function c_asr_gpu_real_create(nsite, nslot, nbr, nsp, species, par, block, block0, r) &
        bind(C, name='asr_gpu_real_create') result(ierr)
      import :: c_ptr, c_int, c_double, c_signed_char
      integer(c_int), value :: nsite, nslot, nsp
      integer(c_int), intent(in) :: nbr(*)
      integer(c_signed_char), intent(in) :: species(*)
      real(c_double), intent(in) :: par(*), block(*), block0(*)
      type(c_ptr), intent(out) :: r
      integer(c_int) :: ierr
    end function
    subroutine c_asr_gpu_real_destroy(r) bind(C, name='asr_gpu_real_destroy')
      import :: c_ptr
      type(c_ptr), value :: r
    end subroutine
    function c_asr_gpu_kpm_moments(r, c, s, nmom, nvec, seed, mu, mu_per, step_ms, total_ms) &
        bind(C, name='asr_gpu_kpm_moments') result(ierr)
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: r, mu_per, step_ms, total_ms
      real(c_double), value :: c, s
      integer(c_int), value :: nmom, nvec
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: mu(*)
      integer(c_int) :: ierr
    end function
  end interface

contains

  !> the engine's message for the last failed call
  function gpu_last_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    call c_f_pointer(c_asr_gpu_last_error(), p, [1024])
    n = 0
    do while (n < 1024)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate (character(len=n) :: msg)
    do n = 1, len(msg); msg(n:n) = p(n); end do
  end function gpu_last_error

  !> error stop with the engine's message on a non-zero status
  subroutine gpu_check(ierr, where)
    integer(c_int), intent(in) :: ierr
    character(*), intent(in) :: where
    if (ierr == 0) return
    print '(a, ": ", a)', where, gpu_last_error()
    error stop 'asr_gpu'
  end subroutine gpu_check

  !> number of visible GPUs (0 without a GPU; the call itself does not fail then)
  integer function gpu_device_count()
    integer(c_int) :: n
    if (c_asr_gpu_device_count(n) /= 0) n = 0
    gpu_device_count = n
  end function gpu_device_count

  !> Device copy of a site-class operator from the tables of augcls_op (src/asr_augcls.f90:45-68) and the unrotated bond
  !> blocks of class_alloy_t: hopr/hopg(nslot, nn) and flipr/flipg(nsmax, nn) with 1-based nodes and operations (hopr = 0:
  !> no link, flipr = 0: no partner), lst(nn) the 0-based local state, ncl(nn) and ntyp(nn) 1-based class and site type
  !> (stype(site(i))), onstart(nn) true on the nodes of the start site (the site augcls_freeze_site freezes), sq(nn),
  !> stab_ptr(nn + 1) / stab(:) 1-based, pi(9, nop) 1-based orbitals and sg(9, nop), block(9, 9, nslot, ntype) =
  !> S_{R, R + bvec(slot)} and block0(9, 9, ntype), nsc(ncls), spc(:, ncls) 1-based species, mt(l + 1, k + 1, s, class) of
  !> class_mt.  The C ABI takes 0-based numbers with -1 for no link and row-major arrays: this is the translation.
  function gpu_op_from_tables(nn, nslot, nop, nsmax, ncls, nsp, ntype, hopr, hopg, flipr, flipg, lst, ncl, ntyp, onstart, &
                              sq, stab_ptr, stab, pi, sg, block, block0, nsc, spc, mt) result(h)
    integer, intent(in) :: nn, nslot, nop, nsmax, ncls, nsp, ntype
    integer, intent(in) :: hopr(nslot, nn), hopg(nslot, nn), flipr(nsmax, nn), flipg(nsmax, nn), lst(nn), ncl(nn), ntyp(nn)
    logical, intent(in) :: onstart(nn)
    real(c_double), intent(in) :: sq(nn)
    integer, intent(in) :: stab_ptr(nn + 1), stab(:), pi(9, nop)
    real(c_double), intent(in) :: sg(9, nop), block(9, 9, nslot, ntype), block0(9, 9, ntype)
    integer, intent(in) :: nsc(ncls), spc(:, :)
    real(c_double), intent(in) :: mt(nsmax, nsmax, nsmax, ncls)
    type(c_ptr) :: h
    integer(c_int), allocatable :: ch(:, :), cg(:, :), cf(:, :), cfg(:, :), cspc(:, :)
    integer(c_signed_char), allocatable :: fz(:)
    real(c_double), allocatable :: cb(:, :, :, :), cb0(:, :, :), cmt(:, :, :, :)
    integer :: k, l, s, kap, t
    call gpu_check(c_asr_gpu_op_create(nn, nslot, nop, nsmax, ncls, nsp, ntype, h), 'asr_gpu_op_create')
    ! links: node n -> n - 1, no link (0) -> -1; operations 1-based -> 0-based (the operation of a missing link is unused)
    allocate (ch(nslot, nn), cg(nslot, nn), cf(nsmax, nn), cfg(nsmax, nn), fz(nn))
    ch = hopr - 1
    cg = merge(hopg - 1, 0, hopr > 0)
    cf = flipr - 1
    cfg = merge(flipg - 1, 0, flipr > 0)
    fz = int(merge(1, 0, onstart), c_signed_char)
    call gpu_check(c_asr_gpu_op_set_tables(h, ch, cg, cf, cfg, lst, ncl - 1, ntyp - 1, fz, sq, stab_ptr - 1, stab - 1, &
                                           size(stab)), 'asr_gpu_op_set_tables')
    call gpu_check(c_asr_gpu_op_set_orbital_ops(h, pi - 1, sg), 'asr_gpu_op_set_orbital_ops')
    ! block[t][k][l][l2] (row-major, l2 fastest) = S(l, l2): the transpose of each column-major 9 x 9 block
    allocate (cb(9, 9, nslot, ntype), cb0(9, 9, ntype))
    do t = 1, ntype
      do k = 1, nslot
        cb(:, :, k, t) = transpose(block(:, :, k, t))
      end do
      cb0(:, :, t) = transpose(block0(:, :, t))
    end do
    call gpu_check(c_asr_gpu_op_set_blocks(h, cb, cb0), 'asr_gpu_op_set_blocks')
    ! spc[class][k] 0-based, -1 beyond nsc; mt[class][l][k][s] (s fastest)
    allocate (cspc(nsmax, ncls), cmt(nsmax, nsmax, nsmax, ncls))
    cspc = -1
    do kap = 1, ncls
      cspc(1:nsc(kap), kap) = spc(1:nsc(kap), kap) - 1
      do l = 1, nsmax
        do k = 1, nsmax
          do s = 1, nsmax
            cmt(s, k, l, kap) = mt(l, k, s, kap)
          end do
        end do
      end do
    end do
    call gpu_check(c_asr_gpu_op_set_classes(h, nsc, cspc, cmt), 'asr_gpu_op_set_classes')
  end function gpu_op_from_tables
end module asr_gpu
