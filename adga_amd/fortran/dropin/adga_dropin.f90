! Shared state of the replacement modules one_particle_quant_module, susc_module and eom_module
! (this directory): one GPU context formed from the reference's parameters_module globals, which are
! the hidden inputs of the module procedures (src/parameters_module.f90:43-73).
!
! A context made before the (q, omega) loop (for get_giw, get_ndmft or get_nfock, which main.F90
! calls during its set-up, src/main.F90:207-271) is formed from whatever globals exist then, with
! placeholders for the ones not yet set (u, u_tilde, n_dmft, the q and EOM k lists); the first loop
! procedure forms it again from the complete set, and it is then kept to the end of the run.
module adga_dropin
  use, intrinsic :: iso_c_binding
  use adga_c
  use parameters_module
  use mpi_org, only: mpi_stop
  implicit none
  private
  public :: dropin_ctx, dropin_check, dropin_release

  type(c_ptr), save :: ctx = c_null_ptr
  logical, save :: ctx_full = .false.

contains

  ! The context; `full` = the caller needs the loop's inputs (u, u_tilde, n_dmft, q_data, k_data_eom).
  function dropin_ctx(full) result(c)
    logical, intent(in) :: full
    type(c_ptr) :: c
    type(adga_dims) :: d
    integer(c_int) :: rc
    integer(c_int), allocatable :: ql(:), kl(:)
    complex(c_double_complex), allocatable :: uu(:,:), ut(:,:), nd(:), gi(:,:)

    if (c_associated(ctx) .and. (ctx_full .or. .not. full)) then
      c = ctx
      return
    endif
    call dropin_release()
    if (ndim > 5) call mpi_stop('adga: the GPU modules take at most 5 orbitals', 1)
    ! the bubble's frequencies: the vertex box, or with the debug option "Bubble" the large range
    ! (config_module.f90:467-477), which the library's option bubble_range = 1 selects
    if ((iwstart /= -iwfmax_small .or. iwstop /= iwfmax_small-1) .and. &
        (iwstart /= -iwmax+iwbmax_small .or. iwstop /= iwmax-iwbmax_small-1)) &
      call mpi_stop('adga: iwstart/iwstop must be the vertex box or the range of the debug option Bubble', 1)
    d%ndim = ndim; d%nkpx = nkpx; d%nkpy = nkpy; d%nkpz = nkpz; d%iwmax = iwmax
    d%iwfmax_small = iwfmax_small; d%iwbmax_small = iwbmax_small
    d%beta = beta; d%mu = mu
    allocate(uu(ndim*ndim, ndim*ndim), ut(ndim*ndim, ndim*ndim), nd(ndim), gi(-iwmax:iwmax-1, ndim))
    uu = 0; ut = 0; nd = 0; gi = 0
    if (allocated(u)) uu = u
    if (allocated(u_tilde)) ut = u_tilde
    if (allocated(n_dmft)) nd = n_dmft
    if (allocated(giw)) gi = giw
    if (full .and. allocated(q_data) .and. allocated(k_data_eom)) then
      allocate(ql(size(q_data)), kl(size(k_data_eom)))
      ql = q_data
      kl = k_data_eom
    else if (full) then
      call mpi_stop('adga: q_data and k_data_eom must be set before the loop', 1)
    else
      allocate(ql(1), kl(1))
      ql = 1
      kl = 1
    endif
    d%nqp = size(ql); d%nkp_eom = size(kl)
    rc = c_adga_ctx_create(d, ql, kl, hk, siw, gi, dc, uu, ut, nd, ctx)
    if (rc /= ADGA_OK) call mpi_stop('adga: GPU context creation failed (adga_ctx_create)', int(rc))
    if (iwstart /= -iwfmax_small) then
      rc = c_adga_set_option(ctx, 'bubble_range'//c_null_char, 1d0)
      if (rc /= ADGA_OK) call mpi_stop('adga: adga_set_option(bubble_range) failed', int(rc))
    endif
    ctx_full = full
    c = ctx
  end function dropin_ctx

  ! mpi_stop with the library's message on a non-zero status, as the reference stops on errors
  subroutine dropin_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(200)
    character(len=200) :: msg
    integer :: i
    integer(c_int) :: dummy
    if (rc == ADGA_OK) return
    msg = ''
    if (c_associated(ctx)) then
      dummy = c_adga_last_error(ctx, buf, 200_c_int)
      do i = 1, 200
        if (buf(i) == c_null_char) exit
        msg(i:i) = buf(i)
      enddo
    endif
    call mpi_stop('adga: '//trim(what)//' failed: '//trim(msg), int(rc))
  end subroutine dropin_check

  subroutine dropin_release()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = c_adga_ctx_destroy(ctx)
    ctx = c_null_ptr
    ctx_full = .false.
  end subroutine dropin_release

end module adga_dropin
