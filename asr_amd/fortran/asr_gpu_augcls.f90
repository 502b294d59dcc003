!> The chain engine behind asr's own types: the device copy of an augcls_op (src/asr_augcls.f90:45-68) with the bond blocks
!> of its class_alloy_t, and G(z) of route edc for a set of energies as edc_green returns it (src/asr_edc.f90:75-164: the
!> coefficients from the engine, the continued fraction by the reference's own cf_species).  Compiled against the
!> reference's modules (asr_kinds, asr_augcls, asr_edc, asr_chain); the part without them is module asr_gpu.
!>
!> Use inside chain_worker (src/asr_class_scf.f90:1093-1157), once per class operator and per batch of tasks:
!>   h = gpu_op_from_augcls(cop(kp)%op, mdl)                    ! after build_augcls_operator, once per SCF run
!>   call gpu_edc_greens(h, ms(pp_, gg), es(pp_, gg), pp_, iorbk(o, kp), zs, nlev, gz(:, o, pp_, gg))
module asr_gpu_augcls
  use, intrinsic :: iso_c_binding
  use asr_kinds, only: dp
  use asr_augcls, only: augcls_op, class_alloy_t
  use asr_edc, only: edn_t
  use asr_chain, only: cf_species, ed_project
  use asr_gpu
  implicit none
  private
  public :: gpu_op_from_augcls, gpu_edc_greens

contains

  !> device copy of op; the unrotated blocks come from mdl (augcls_set_blocks), the site types from op%stype, the start
  !> site is the site of node 1 (the start node of build_augcls_operator)
  function gpu_op_from_augcls(op, mdl) result(h)
    type(augcls_op), intent(in) :: op
    type(class_alloy_t), intent(in) :: mdl
    type(c_ptr) :: h
    integer, allocatable :: ntyp(:)
    real(dp), allocatable :: b0(:, :, :)
    allocate (ntyp(op%nn))
    if (allocated(op%stype)) then
      ntyp = op%stype(op%site)
    else
      ntyp = 1
    end if
    allocate (b0(9, 9, op%ntype)); b0 = 0.0_dp
    if (allocated(mdl%block0)) b0 = mdl%block0(:, :, 1:op%ntype)
    h = gpu_op_from_tables(op%nn, op%nslot, op%nop, op%nsmax, op%ncls, op%nsp, op%ntype, op%hopr, op%hopg, op%flipr, &
                           op%flipg, op%lst, op%ncl, ntyp, op%site == op%site(1), op%sq, op%stab_ptr, op%stab, op%pi, op%sg, &
                           mdl%block(:, :, 1:op%nslot, 1:op%ntype), b0, op%nsc, op%spc, op%mt)
  end function gpu_op_from_augcls

  !> G_c(z) of the start orbital l0 (1-based) at every energy of zs, central species `central` (the frozen start site),
  !> species parameters of edc_setup (mdl%c, %d, %o, %e) and e_s(z) = z pa + pb of ed (edc_green_blk, src/asr_edc.f90:
  !> 253-258), nlev levels, the projection as ed_project; one batched call of asr_gpu_chains_edc, then cf_species per chain
  subroutine gpu_edc_greens(h, mdl, ed, central, l0, zs, nlev, g)
    type(c_ptr), intent(in) :: h
    type(class_alloy_t), intent(in) :: mdl
    type(edn_t), intent(in) :: ed
    integer, intent(in) :: central, l0, nlev
    complex(dp), intent(in) :: zs(:)
    complex(dp), intent(out) :: g(:)
    integer :: nz, j, s, l
    integer(c_int), allocatable :: so(:), fz(:), nl(:), mu(:)
    real(c_double), allocatable :: par(:, :, :, :), pab(:, :, :, :)
    complex(c_double_complex), allocatable :: a(:, :), b(:, :)
    complex(dp) :: f
    complex(dp), allocatable :: cc(:)
    real(c_double) :: ms
    nz = size(zs)
    allocate (so(nz), fz(nz), nl(nz), mu(nz), par(9, mdl%nsp, 4, nz), pab(9, mdl%nsp, 2, nz), a(nlev, nz), b(nlev, nz), &
              cc(nlev))
    so = l0 - 1; fz = central - 1; nl = nlev
    do j = 1, nz
      par(:, :, 1, j) = mdl%c; par(:, :, 2, j) = mdl%d; par(:, :, 3, j) = mdl%o; par(:, :, 4, j) = mdl%e
      do s = 1, mdl%nsp
        do l = 1, 9
          pab(l, s, 1, j) = 1.0_dp - ed%dc(l) / ed%del(l, s)
          pab(l, s, 2, j) = ed%dc(l) * ed%enu(l, s) / ed%del(l, s)
        end do
      end do
    end do
    call gpu_check(c_asr_gpu_chains_edc(h, nz, so, par, pab, fz, c_null_ptr, zs, nl, merge(1, 0, ed_project), nlev, a, b, &
                                        mu, 0, ms), 'asr_gpu_chains_edc')
    cc = (0.0_dp, 0.0_dp)
    do j = 1, nz
      call cf_species(a(:, j), b(:, j), cc, mu(j), zs(j), g(j), f)
    end do
  end subroutine gpu_edc_greens
end module asr_gpu_augcls
