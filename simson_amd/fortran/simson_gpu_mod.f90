! ISO_C_BINDING interface to the MI355X Simson step library
! (include/simson_gpu.h).  Arrays are passed in the reference main-storage
! layout ur(memnx,memny,memnz,memnxyz) of bla/bla.f:67.
module simson_gpu_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SIMSON_FP64 = 0

  type, bind(c) :: simson_params_t
     integer(c_int) :: nx, ny, nz
     integer(c_int) :: precision
     integer(c_int) :: ibc
     integer(c_int) :: cflux
     real(c_double) :: re
     real(c_double) :: xl, zl
     real(c_double) :: rot
     real(c_double) :: retau
  end type simson_params_t

  ! ADM-RT LES parameters (sgs.i, rparamles.f)
  type, bind(c) :: simson_fringe_t
     real(c_double) :: fmax
     real(c_double) :: fstart
     real(c_double) :: fend
     real(c_double) :: frise
     real(c_double) :: ffall
     real(c_double) :: xsc
  end type simson_fringe_t

  type, bind(c) :: simson_wall_t
     real(c_double) :: amp, damp, xstart, xend, xrise, xfall, zbet, tomeg, xsc, zsc
  end type simson_wall_t

  ! spatial boundary-layer base flow (cbflow.f), units of the .u file
  type, bind(c) :: simson_baseflow_t
     integer(c_int) :: nx, ny, fltype
     real(c_double) :: xl, re, dstar, dstar2, bstart, blength, rlam, spanv, xsc, u0low, w0low
  end type simson_baseflow_t

  ! localized volume force of bla.i loctyp = 3 (locf.f)
  type, bind(c) :: simson_locf_t
     real(c_double) :: ampx, ampy, ampz, xscale, xloc0, yscale, xtype, tomeg, xsc
  end type simson_locf_t

  type, bind(c) :: simson_les_t
     integer(c_int) :: iles
     integer(c_int) :: iord
     real(c_double) :: cutoff
     real(c_double) :: chi
     real(c_double) :: cs
     integer(c_int) :: ineg
  end type simson_les_t

  interface
     integer(c_int) function simson_gpu_init(p, g) bind(c, name='simson_gpu_init')
       import :: c_int, c_ptr, simson_params_t
       type(simson_params_t), intent(in) :: p
       type(c_ptr), intent(out) :: g
     end function simson_gpu_init

     integer(c_int) function simson_gpu_upload(g, ur, ui, nslots, puw, px) &
          bind(c, name='simson_gpu_upload')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(in) :: ur(*), ui(*)
       integer(c_int), value :: nslots
       type(c_ptr), value :: puw
       real(c_double), value :: px
     end function simson_gpu_upload

     integer(c_int) function simson_gpu_prhs(g, an) bind(c, name='simson_gpu_prhs')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), value :: an
     end function simson_gpu_prhs

     integer(c_int) function simson_gpu_nonlin(g, do_cfl, cfl_out) &
          bind(c, name='simson_gpu_nonlin')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       integer(c_int), value :: do_cfl
       real(c_double), intent(out) :: cfl_out
     end function simson_gpu_nonlin

     integer(c_int) function simson_gpu_cfl(g, cfl) bind(c, name='simson_gpu_cfl')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(out) :: cfl
     end function simson_gpu_cfl

     integer(c_int) function simson_gpu_linear(g, an, bn, anp1, bnp1, px_out) &
          bind(c, name='simson_gpu_linear')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), value :: an, bn, anp1, bnp1
       type(c_ptr), value :: px_out
     end function simson_gpu_linear

     integer(c_int) function simson_gpu_download(g, ur, ui, nslots, puw, px) &
          bind(c, name='simson_gpu_download')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(out) :: ur(*), ui(*)
       integer(c_int), value :: nslots
       type(c_ptr), value :: puw
       real(c_double), intent(out) :: px
     end function simson_gpu_download

     integer(c_int) function simson_gpu_amp(g, amp) bind(c, name='simson_gpu_amp')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(out) :: amp(*)
     end function simson_gpu_amp

     integer(c_int) function simson_gpu_pressure(g, pr, pim) &
          bind(c, name='simson_gpu_pressure')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(out) :: pr(*), pim(*)
     end function simson_gpu_pressure

     integer(c_int) function simson_gpu_xys(g, dtn, xs, zs, nxys, xys) &
          bind(c, name='simson_gpu_xys')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), value :: dtn, xs, zs
       integer(c_int), value :: nxys
       real(c_double), intent(inout) :: xys(*)
     end function simson_gpu_xys

     integer(c_int) function simson_gpu_fringe(g, f, base) bind(c, name='simson_gpu_fringe')
       import :: c_int, c_ptr, c_double, simson_fringe_t
       type(c_ptr), value :: g
       type(simson_fringe_t), intent(in) :: f
       real(c_double), intent(in) :: base(*)
     end function simson_gpu_fringe

     integer(c_int) function simson_gpu_free_stream(g, bjr, bji, du0upp) &
          bind(c, name='simson_gpu_free_stream')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(in) :: bjr(*), bji(*)
       real(c_double), value :: du0upp
     end function simson_gpu_free_stream

     integer(c_int) function simson_base_flow(b, tab, ntab, base, bjr, bji, du0upp) &
          bind(c, name='simson_base_flow')
       import :: c_int, c_double, simson_baseflow_t
       type(simson_baseflow_t), intent(in) :: b
       real(c_double), intent(in) :: tab(*)
       integer(c_int), value :: ntab
       real(c_double), intent(out) :: base(*), bjr(*), bji(*)
       real(c_double), intent(out) :: du0upp
     end function simson_base_flow

     integer(c_int) function simson_parallel_force(b, tab, ntab, cdev, t, force) &
          bind(c, name='simson_parallel_force')
       import :: c_int, c_double, simson_baseflow_t
       type(simson_baseflow_t), intent(in) :: b
       real(c_double), intent(in) :: tab(*)
       integer(c_int), value :: ntab
       real(c_double), value :: cdev, t
       real(c_double), intent(out) :: force(*)
     end function simson_parallel_force

     integer(c_int) function simson_gpu_mean_force(g, force) bind(c, name='simson_gpu_mean_force')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), intent(in) :: force(*)
     end function simson_gpu_mean_force

     integer(c_int) function simson_gpu_local_force(g, lf) bind(c, name='simson_gpu_local_force')
       import :: c_int, c_ptr, simson_locf_t
       type(c_ptr), value :: g
       type(simson_locf_t), intent(in) :: lf
     end function simson_gpu_local_force

     integer(c_int) function simson_gpu_wall(g, w) bind(c, name='simson_gpu_wall')
       import :: c_int, c_ptr, simson_wall_t
       type(c_ptr), value :: g
       type(simson_wall_t), intent(in) :: w
     end function simson_gpu_wall

     integer(c_int) function simson_gpu_set_time(g, t) bind(c, name='simson_gpu_set_time')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: g
       real(c_double), value :: t
     end function simson_gpu_set_time

     integer(c_int) function simson_gpu_les(g, les) bind(c, name='simson_gpu_les')
       import :: c_int, c_ptr, simson_les_t
       type(c_ptr), value :: g
       type(simson_les_t), intent(in) :: les
     end function simson_gpu_les

     subroutine simson_gpu_finalize(g) bind(c, name='simson_gpu_finalize')
       import :: c_ptr
       type(c_ptr), value :: g
     end subroutine simson_gpu_finalize
  end interface
end module simson_gpu_mod
