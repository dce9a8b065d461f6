! ISO_C_BINDING interfaces of the x2gpu C ABI (include/x2gpu.h).  This is the
! only place where the Fortran driver touches the device: every call passes
! plain pointers and sizes and returns an integer status (0 = success).
module x2f_capi
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: X2G_OK = 0

  type, bind(c) :: x2g_relax_item
     type(c_ptr) :: fun_d = c_null_ptr
     type(c_ptr) :: lhs_d = c_null_ptr
     type(c_ptr) :: rhs_d = c_null_ptr
     type(c_ptr) :: tail_fock1_d = c_null_ptr
     real(c_double) :: omega = 0.0_c_double
     integer(c_int) :: isym = 1
     integer(c_int) :: bkind = 1
     integer(c_int) :: prec = 0
     integer(c_int) :: grid_index = 0
     type(c_ptr) :: rhs_f1_d = c_null_ptr   ! rhs = f1*f2*f3 formed by the residual (optional)
     type(c_ptr) :: rhs_f2_d = c_null_ptr
     type(c_ptr) :: rhs_f3_d = c_null_ptr
     integer(c_int) :: method = 0            ! 0 MCSOR, 1 SOR, 2 backward SOR (x2g_relax_ordered)
     integer(c_int) :: ordering = 0          ! X2G_ORDER_*: 0 col-wise, 1 middle, .. 9 middle2
  end type x2g_relax_item

  ! the device SCF control record (x2g_scf_ctrl); orbital i is index i
  integer, parameter :: X2G_CTRL_MAXO = 16
  type, bind(c) :: x2g_scf_ctrl
     real(c_double) :: ethld = 0, nthld = 0, thren4init = 0, recalc_mm = 0, inc_thld = 0
     real(c_double) :: ee(X2G_CTRL_MAXO) = 0, eeold(X2G_CTRL_MAXO) = 0
     real(c_double) :: deltaee(X2G_CTRL_MAXO) = 0, orbnorm(X2G_CTRL_MAXO) = 0
     real(c_double) :: eeconv = 0, eeconvprev = 0, dnmaxprev = 0, deltaeemax = 0, deltanormmax = 0
     integer(c_int) :: norb = 0, maxscf = 0, nscf2skip = 0, nenlast = 0, nnolast = 0, nscfextra = 0
     integer(c_int) :: maxresets = 0, polish = 0
     integer(c_int) :: ifixorb(X2G_CTRL_MAXO) = 0, itouch(X2G_CTRL_MAXO) = 0, mm_touch(X2G_CTRL_MAXO) = 0
     integer(c_int) :: iscf = 0, nresets = 0, nresets4no = 0, norbenergythld = 0, norbnormthld = 0
     integer(c_int) :: noenergydec = 0, nonormdec = 0
     integer(c_int) :: lenergythld = 0, lnormthld = 0, lfixorb = 0, lfixorb4init = 0
     integer(c_int) :: mm_due = 0, prec = 0, polish_end = 0, stopcode = 0, stop_flag = 0
  end type x2g_scf_ctrl

  type, bind(c) :: x2g_term
     real(c_double) :: c = 0.0_c_double
     type(c_ptr) :: a_d = c_null_ptr
     type(c_ptr) :: b_d = c_null_ptr
     type(c_ptr) :: cc_d = c_null_ptr
  end type x2g_term

  type, bind(c) :: x2g_dot
     type(c_ptr) :: w_d = c_null_ptr
     type(c_ptr) :: a_d = c_null_ptr
     type(c_ptr) :: b_d = c_null_ptr
     type(c_ptr) :: cc_d = c_null_ptr
  end type x2g_dot

  integer, parameter :: X2G_MAX_ORTHO = 16

  type, bind(c) :: x2g_lcao
     integer(c_int) :: m = 0, n1 = 1, l1 = 0, n2 = 1, l2 = 0
     real(c_double) :: c1 = 0.0_c_double, za = 0.0_c_double
     real(c_double) :: c2 = 0.0_c_double, zb = 0.0_c_double
  end type x2g_lcao

  type, bind(c) :: x2g_lincomb_op
     type(c_ptr) :: out_d = c_null_ptr
     type(c_ptr) :: base_d = c_null_ptr
     real(c_double) :: base_scale = 0.0_c_double
     integer(c_int) :: first_term = 0
     integer(c_int) :: nterms = 0
     integer(c_int) :: grid_index = 0
  end type x2g_lincomb_op

  type, bind(c) :: x2g_normortho_item
     type(c_ptr) :: psi_d = c_null_ptr
     type(c_ptr) :: others_d(X2G_MAX_ORTHO) = c_null_ptr
     integer(c_int) :: nothers = 0
     integer(c_int) :: ihomo = 0
     integer(c_int) :: normalise = 1
     integer(c_int) :: grid_index = 0
     type(c_ptr) :: norm_out_d = c_null_ptr
  end type x2g_normortho_item

  type, bind(c) :: x2g_eahf_item
     type(c_ptr) :: psi_d = c_null_ptr
     type(c_ptr) :: cpot_d = c_null_ptr
     type(c_ptr) :: xpot_d = c_null_ptr
     integer(c_int) :: isym = 1
     integer(c_int) :: m = 0
     integer(c_int) :: grid_index = 0
     type(c_ptr) :: ee_out_d = c_null_ptr
  end type x2g_eahf_item

  type, bind(c) :: x2g_moment_item
     type(c_ptr) :: psi1_d = c_null_ptr
     type(c_ptr) :: psi2_d = c_null_ptr
     integer(c_int) :: dm = 0
     integer(c_int) :: grid_index = 0
  end type x2g_moment_item

  type, bind(c) :: x2g_tail_item
     type(c_ptr) :: pot_d = c_null_ptr
     integer(c_int) :: kind = 0
     integer(c_int) :: dm = 0
     integer(c_int) :: grid_index = 0
     real(c_double) :: mom(8) = 0.0_c_double
  end type x2g_tail_item

  interface
     function x2g_last_error() bind(c, name='x2g_last_error')
       import :: c_ptr
       type(c_ptr) :: x2g_last_error
     end function x2g_last_error

     integer(c_int) function x2g_ctx_create(device, out) bind(c, name='x2g_ctx_create')
       import :: c_int, c_ptr
       integer(c_int), value :: device
       type(c_ptr) :: out
     end function x2g_ctx_create

     integer(c_int) function x2g_ctx_destroy(ctx) bind(c, name='x2g_ctx_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function x2g_ctx_destroy

     integer(c_int) function x2g_ctx_sync(ctx) bind(c, name='x2g_ctx_sync')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function x2g_ctx_sync

     integer(c_int) function x2g_malloc(ctx, bytes, dptr) bind(c, name='x2g_malloc')
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: ctx
       integer(c_size_t), value :: bytes
       type(c_ptr) :: dptr
     end function x2g_malloc

     integer(c_int) function x2g_free(ctx, dptr) bind(c, name='x2g_free')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, dptr
     end function x2g_free

     integer(c_int) function x2g_h2d(ctx, dst, src, bytes) bind(c, name='x2g_h2d')
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: ctx, dst, src
       integer(c_size_t), value :: bytes
     end function x2g_h2d

     integer(c_int) function x2g_d2h(ctx, dst, src, bytes) bind(c, name='x2g_d2h')
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: ctx, dst, src
       integer(c_size_t), value :: bytes
     end function x2g_d2h

     integer(c_int) function x2g_d2d(ctx, dst, src, bytes) bind(c, name='x2g_d2d')
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: ctx, dst, src
       integer(c_size_t), value :: bytes
     end function x2g_d2d

     integer(c_int) function x2g_fill(ctx, dst, n, value) bind(c, name='x2g_fill')
       import :: c_int, c_ptr, c_size_t, c_double
       type(c_ptr), value :: ctx, dst
       integer(c_size_t), value :: n
       real(c_double), value :: value
     end function x2g_fill

     integer(c_int) function x2g_guess_orbital(g, orb, lc) bind(c, name='x2g_guess_orbital')
       import :: c_int, c_ptr, x2g_lcao
       type(c_ptr), value :: g, orb
       type(x2g_lcao), intent(in) :: lc
     end function x2g_guess_orbital

     integer(c_int) function x2g_guess_coulomb(g, pot, c1, c2, two_centres) bind(c, name='x2g_guess_coulomb')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g, pot
       real(c_double), value :: c1, c2
       integer(c_int), value :: two_centres
     end function x2g_guess_coulomb

     integer(c_int) function x2g_guess_exchange(g, pot, c1, c2) bind(c, name='x2g_guess_exchange')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g, pot
       real(c_double), value :: c1, c2
     end function x2g_guess_exchange

     integer(c_int) function x2g_grid_create(ctx, nni, nmu, coef, suppl, vec, out) &
          bind(c, name='x2g_grid_create')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nni, nmu
       real(c_double), intent(in) :: coef(*), suppl(*), vec(*)
       type(c_ptr) :: out
     end function x2g_grid_create

     integer(c_int) function x2g_grid_destroy(g) bind(c, name='x2g_grid_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: g
     end function x2g_grid_destroy

     function x2g_grid_suppl(g, k) bind(c, name='x2g_grid_suppl')
       import :: c_int, c_ptr
       type(c_ptr), value :: g
       integer(c_int), value :: k
       type(c_ptr) :: x2g_grid_suppl
     end function x2g_grid_suppl

     integer(c_int) function x2g_relax_batch(ctx, grids, items, n, maxsor1, maxsor2) &
          bind(c, name='x2g_relax_batch')
       import :: c_int, c_ptr, x2g_relax_item
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: grids(*)
       type(x2g_relax_item), intent(in) :: items(*)
       integer(c_int), value :: n, maxsor1, maxsor2
     end function x2g_relax_batch

     integer(c_long_long) function x2g_relax_workspace_bytes(nni, nmu, prec) &
          bind(c, name='x2g_relax_workspace_bytes')
       import :: c_int, c_long_long
       integer(c_int), value :: nni, nmu, prec
     end function x2g_relax_workspace_bytes

     integer(c_int) function x2g_lincomb(g, out, base, base_scale, terms, nterms) &
          bind(c, name='x2g_lincomb')
       import :: c_int, c_ptr, c_double, x2g_term
       type(c_ptr), value :: g, out, base
       real(c_double), value :: base_scale
       type(x2g_term), intent(in) :: terms(*)
       integer(c_int), value :: nterms
     end function x2g_lincomb

     integer(c_int) function x2g_wdot(g, terms, nterms, out) bind(c, name='x2g_wdot')
       import :: c_int, c_ptr, c_double, x2g_dot
       type(c_ptr), value :: g
       type(x2g_dot), intent(in) :: terms(*)
       integer(c_int), value :: nterms
       real(c_double), intent(out) :: out(*)
     end function x2g_wdot

     integer(c_int) function x2g_oneel(g, psi1, psi2, isym, m, out) bind(c, name='x2g_oneel')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g, psi1, psi2
       integer(c_int), value :: isym, m
       real(c_double), intent(out) :: out(3)
     end function x2g_oneel

     integer(c_int) function x2g_moments(g, psi1, psi2, dm, mpole, out) bind(c, name='x2g_moments')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g, psi1, psi2
       integer(c_int), value :: dm, mpole
       real(c_double), intent(out) :: out(*)
     end function x2g_moments

     integer(c_int) function x2g_pot_tail(g, pot, kind, dm, mpole, mom) bind(c, name='x2g_pot_tail')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g, pot
       integer(c_int), value :: kind, dm, mpole
       real(c_double), intent(in) :: mom(*)
     end function x2g_pot_tail

     integer(c_int) function x2g_lincomb_batch(ctx, grids, ops, nops, terms, nterms) &
          bind(c, name='x2g_lincomb_batch')
       import :: c_int, c_ptr, x2g_lincomb_op, x2g_term
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: grids(*)
       type(x2g_lincomb_op), intent(in) :: ops(*)
       integer(c_int), value :: nops
       type(x2g_term), intent(in) :: terms(*)
       integer(c_int), value :: nterms
     end function x2g_lincomb_batch

     integer(c_int) function x2g_normortho_batch(ctx, grids, items, n) bind(c, name='x2g_normortho_batch')
       import :: c_int, c_ptr, x2g_normortho_item
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: grids(*)
       type(x2g_normortho_item), intent(in) :: items(*)
       integer(c_int), value :: n
     end function x2g_normortho_batch

     integer(c_int) function x2g_eahf_batch(ctx, grids, items, n) bind(c, name='x2g_eahf_batch')
       import :: c_int, c_ptr, x2g_eahf_item
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: grids(*)
       type(x2g_eahf_item), intent(in) :: items(*)
       integer(c_int), value :: n
     end function x2g_eahf_batch

     integer(c_int) function x2g_moments_batch(ctx, grids, items, n, mpole, out) bind(c, name='x2g_moments_batch')
       import :: c_int, c_ptr, c_double, x2g_moment_item
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: grids(*)
       type(x2g_moment_item), intent(in) :: items(*)
       integer(c_int), value :: n, mpole
       real(c_double), intent(out) :: out(*)
     end function x2g_moments_batch

     integer(c_int) function x2g_pot_tail_batch(ctx, grids, items, n, mpole) bind(c, name='x2g_pot_tail_batch')
       import :: c_int, c_ptr, x2g_tail_item
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: grids(*)
       type(x2g_tail_item), intent(in) :: items(*)
       integer(c_int), value :: n, mpole
     end function x2g_pot_tail_batch

     integer(c_int) function x2g_gather(ctx, src, counts, n, out) bind(c, name='x2g_gather')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: src(*)
       integer(c_int), intent(in) :: counts(*)
       integer(c_int), value :: n
       real(c_double), intent(out) :: out(*)
     end function x2g_gather

     integer(c_int) function x2g_scf_control(ctx, ctrl, rec, n) bind(c, name='x2g_scf_control')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), intent(in) :: ctrl(*), rec(*)
       integer(c_int), value :: n
     end function x2g_scf_control

     integer(c_int) function x2g_scf_ctrl_bytes() bind(c, name='x2g_scf_ctrl_bytes')
       import :: c_int
     end function x2g_scf_ctrl_bytes

     integer(c_int) function x2g_setci(g, psi, ihsym) bind(c, name='x2g_setci')
       import :: c_int, c_ptr
       type(c_ptr), value :: g, psi
       integer(c_int), value :: ihsym
     end function x2g_setci
  end interface

contains

  ! Stop the job with the library's message when a C-ABI call failed.
  subroutine x2check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    character(len=512) :: txt
    if (rc == X2G_OK) return
    call c_f_pointer(x2g_last_error(), msg, [512])
    txt = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       txt(i:i) = msg(i)
    enddo
    write(0, '("x2dhf_amd: ",a," failed (",i0,"): ",a)') where, rc, trim(txt)
    error stop 2
  end subroutine x2check

  ! Device pointer offset by `k` doubles.
  function dptr(p, k) result(q)
    type(c_ptr), intent(in) :: p
    integer(c_size_t), intent(in) :: k
    type(c_ptr) :: q
    integer(c_intptr_t) :: a
    a = transfer(p, a) + 8_c_intptr_t*int(k, c_intptr_t)
    q = transfer(a, q)
  end function dptr
end module x2f_capi
