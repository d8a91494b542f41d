! ermgpu.F90 -- ISO_C_BINDING interface of libermgpu.so (include/ermgpu.h)
! for a Fortran host such as ERmod's engproc.  Every entry point of the
! C-ABI is declared here with its C types; indices are 0-based, matrices
! column-major (the Fortran layout of engmain's cell and ecorr).
module ermgpu
  use iso_c_binding, only: c_int32_t, c_int64_t, c_double, c_ptr, c_char, c_null_char, &
                           c_f_pointer, c_null_ptr, c_associated
  implicit none
  private :: c_int32_t, c_int64_t, c_double, c_char, c_null_char, c_f_pointer

  integer(c_int32_t), parameter :: ERM_OK = 0, ERM_E_ENG_MIN = 1, ERM_E_ENG_SFT = 2, &
       ERM_E_MBIN = 3, ERM_E_ENG_ECD = 4, ERM_E_RCP_CNS = 5, ERM_E_NONCUBOID = 6, &
       ERM_E_ENG_BUG = 7, ERM_E_ENG_CNS = 8, ERM_E_ENG_SIZ = 9, ERM_E_ARG = 10, &
       ERM_E_UNSUPPORTED = 11, ERM_E_DEVICE = 12, ERM_E_EOF = 13, ERM_E_INS_LIMIT = 14
  integer(c_int32_t), parameter :: ERM_FP64 = 0, ERM_FP32 = 1

  type, bind(C) :: erm_params
     integer(c_int32_t) :: slttype, estype, ljswitch, splodr
     integer(c_int32_t) :: ms(3)
     integer(c_int32_t) :: corrcal, selfcal, wgtslf, force_calculation
     real(c_double) :: elecut, lwljcut, upljcut, screen, temp, block_threshold
     integer(c_int32_t) :: precision, max_batch, device, pme_path
     ! engmain's cltype (0 bare Coulomb, 2 PME, 3 PPPM) and boxshp
     integer(c_int32_t) :: cltype, boxshp
  end type erm_params

  type, bind(C) :: erm_hist
     real(c_double) :: ecdmin, ecfmns, ecdcen, eccore, ecdmax
     real(c_double) :: eclbin, ecfbin, ec0bin, finfac
     integer(c_int32_t) :: pecore
  end type erm_hist

  interface
     integer(c_int32_t) function erm_build_edges(h, self_energy, ercrd, cap, pemax, pesoft) &
          bind(C, name="erm_build_edges")
       import :: erm_hist, c_int32_t, c_double
       type(erm_hist), intent(in) :: h
       integer(c_int32_t), value :: self_energy, cap
       real(c_double), intent(out) :: ercrd(*)
       integer(c_int32_t), intent(out) :: pemax, pesoft
     end function erm_build_edges

     integer(c_int64_t) function erm_hash_double(v, n) bind(C, name="erm_hash_double")
       import :: c_int64_t, c_int32_t, c_double
       real(c_double), intent(in) :: v(*)
       integer(c_int32_t), value :: n
     end function erm_hash_double

     integer(c_int32_t) function erm_create(params, ctx) bind(C, name="erm_create")
       import :: erm_params, c_int32_t, c_ptr
       type(erm_params), intent(in) :: params
       type(c_ptr), intent(out) :: ctx
     end function erm_create

     subroutine erm_destroy(ctx) bind(C, name="erm_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine erm_destroy

     type(c_ptr) function erm_last_error_c(ctx) bind(C, name="erm_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function erm_last_error_c

     integer(c_int32_t) function erm_set_topology(ctx, natom, nmol, mol_begin, charge, ljtype, ntype, &
          lj_eps, lj_sig2, sluvid, uvspec, mol_charge, sitemass, bfcoord) bind(C, name="erm_set_topology")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: natom, nmol, ntype
       integer(c_int32_t), intent(in) :: mol_begin(*), ljtype(*), sluvid(*), uvspec(*)
       real(c_double), intent(in) :: charge(*), lj_eps(*), lj_sig2(*), mol_charge(*), sitemass(*), bfcoord(*)
     end function erm_set_topology

     integer(c_int32_t) function erm_set_bins(ctx, numslv, uvmax, uvsoft, uvcrd, esmax, escrd) &
          bind(C, name="erm_set_bins")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: numslv, esmax
       integer(c_int32_t), intent(in) :: uvmax(*), uvsoft(*)
       real(c_double), intent(in) :: uvcrd(*), escrd(*)
     end function erm_set_bins

     integer(c_int32_t) function erm_engclear(ctx) bind(C, name="erm_engclear")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
     end function erm_engclear

     integer(c_int32_t) function erm_set_frame(ctx, sitepos, cell, weight) bind(C, name="erm_set_frame")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: sitepos(*), cell(3, 3)
       real(c_double), value :: weight
     end function erm_set_frame

     integer(c_int32_t) function erm_set_frame_device(ctx, d_sitepos, cell, weight) &
          bind(C, name="erm_set_frame_device")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx, d_sitepos
       real(c_double), intent(in) :: cell(3, 3)
       real(c_double), value :: weight
     end function erm_set_frame_device

     integer(c_int32_t) function erm_get_uv_energy(ctx, solute, n, solute_pos, uvengy) &
          bind(C, name="erm_get_uv_energy")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: solute, n
       real(c_double), intent(in) :: solute_pos(*)
       real(c_double), intent(out) :: uvengy(*)
     end function erm_get_uv_energy

     integer(c_int32_t) function erm_realcal_proc(ctx, solute, solute_pos, ereal) &
          bind(C, name="erm_realcal_proc")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: solute
       real(c_double), intent(in) :: solute_pos(*)
       real(c_double), intent(out) :: ereal(*)
     end function erm_realcal_proc

     integer(c_int32_t) function erm_realcal_self(ctx, solute, solute_pos, self_real) &
          bind(C, name="erm_realcal_self")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: solute
       real(c_double), intent(in) :: solute_pos(*)
       real(c_double), intent(out) :: self_real
     end function erm_realcal_self

     integer(c_int32_t) function erm_recpcal(ctx, solute, solute_pos, erecp, self_recp) &
          bind(C, name="erm_recpcal")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: solute
       real(c_double), intent(in) :: solute_pos(*)
       real(c_double), intent(out) :: erecp(*), self_recp
     end function erm_recpcal

     integer(c_int32_t) function erm_engconst(ctx, n, solutes, solute_pos, weights, uvengy) &
          bind(C, name="erm_engconst")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       type(c_ptr), value :: solutes, solute_pos, weights, uvengy
     end function erm_engconst

     integer(c_int32_t) function erm_insert_random(ctx, n, seed, first_index, uvengy) &
          bind(C, name="erm_insert_random")
       import :: c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       integer(c_int64_t), value :: seed, first_index
       type(c_ptr), value :: uvengy
     end function erm_insert_random

     integer(c_int32_t) function erm_insert_random_multi(ctxs, nctx, n, seed, first_index) &
          bind(C, name="erm_insert_random_multi")
       import :: c_ptr, c_int32_t, c_int64_t
       type(c_ptr), intent(in) :: ctxs(*)
       integer(c_int32_t), value :: nctx
       integer(c_int32_t), intent(in) :: n(*)
       integer(c_int64_t), intent(in) :: seed(*), first_index(*)
     end function erm_insert_random_multi

     integer(c_int32_t) function erm_update_histogram(ctx, n, solutes, weights, uvengy) &
          bind(C, name="erm_update_histogram")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       type(c_ptr), value :: solutes, weights, uvengy
     end function erm_update_histogram

     integer(c_int32_t) function erm_engstore(ctx) bind(C, name="erm_engstore")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
     end function erm_engstore

     integer(c_int32_t) function erm_fetch(ctx, edens, ecorr, eself, minuv, maxuv, slnuv, scalars) &
          bind(C, name="erm_fetch")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx, edens, ecorr, eself, minuv, maxuv, slnuv, scalars
     end function erm_fetch

     integer(c_int32_t) function erm_sizes(ctx, slvmax, ermax) bind(C, name="erm_sizes")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: slvmax, ermax
     end function erm_sizes

     integer(c_int32_t) function erm_stage_times(ctx, ms10) bind(C, name="erm_stage_times")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms10(10)
     end function erm_stage_times

     integer(c_int32_t) function erm_set_tuning(ctx, max_batch, pme_path, use_graph, slab_planes, &
          group_planes, slab_threads, column_threads, fold_y) bind(C, name="erm_set_tuning")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: max_batch, pme_path, use_graph, slab_planes, group_planes, slab_threads, &
                                    column_threads, fold_y
     end function erm_set_tuning

     ! real-space kernel: 0 compacted tiles, 1 one thread per candidate molecule
     integer(c_int32_t) function erm_set_pair_path(ctx, path) bind(C, name="erm_set_pair_path")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: path
     end function erm_set_pair_path

     ! insorigin / insposition / insorient / lwreg / upreg / hostspec of
     ! instslt (insertion.F90:103-366), the reference's numbering
     integer(c_int32_t) function erm_set_insertion(ctx, insorigin, insposition, insorient, lwreg, upreg, &
          nhost, hostspec) bind(C, name="erm_set_insertion")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: insorigin, insposition, insorient, nhost
       real(c_double), value :: lwreg, upreg
       integer(c_int32_t), intent(in) :: hostspec(*)
     end function erm_set_insertion

     ! random placement of flexible-solute bodies (SltConf) with weights
     integer(c_int32_t) function erm_insert_flexible(ctx, n, confs, weights, seed, first_index, uvengy) &
          bind(C, name="erm_insert_flexible")
       import :: c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n
       type(c_ptr), value :: confs, weights, uvengy
       integer(c_int64_t), value :: seed, first_index
     end function erm_insert_flexible

     ! load_refstructure's arrays for the REFSTR insertion modes
     integer(c_int32_t) function erm_set_refstructure(ctx, nhost, host_atoms, host_crd, host_weight, &
          slt_crd, slt_weight) bind(C, name="erm_set_refstructure")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: nhost
       integer(c_int32_t), intent(in) :: host_atoms(*)
       real(c_double), intent(in) :: host_crd(*), host_weight(*), slt_crd(*), slt_weight(*)
     end function erm_set_refstructure

     integer(c_int32_t) function erm_get_tuning(ctx, out10) bind(C, name="erm_get_tuning")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: out10(10)
     end function erm_get_tuning

     integer(c_int32_t) function erm_autotune(ctx, n_trial, seed, best_ms) bind(C, name="erm_autotune")
       import :: c_ptr, c_int32_t, c_int64_t, c_double
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n_trial
       integer(c_int64_t), value :: seed
       real(c_double), intent(out) :: best_ms
     end function erm_autotune

     integer(c_int32_t) function erm_solute_gradient(ctx, n, solutes, solute_pos, seed, first_index, &
          weights, cmbrule, q, eps, sigma, acc) bind(C, name="erm_solute_gradient")
       import :: c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: n, cmbrule
       integer(c_int64_t), value :: seed, first_index
       type(c_ptr), value :: solutes, solute_pos, weights, q, eps, sigma, acc
     end function erm_solute_gradient

     ! the intermediate-state method (extensions/Intermediate/engproc.F90)
     integer(c_int32_t) function erm_set_intermediate(ctx, zerodns) bind(C, name="erm_set_intermediate")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(in) :: zerodns(*)
     end function erm_set_intermediate

     integer(c_int32_t) function erm_fetch_intermediate(ctx, edintm, uvintm, scal) &
          bind(C, name="erm_fetch_intermediate")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: edintm(*), uvintm(*), scal(3)
     end function erm_fetch_intermediate

     ! frame ingestion (getconf_parallel, setconf.F90:860-996; dcd.f90)
     integer(c_int32_t) function erm_traj_open_c(ctx, dcd_path, natom_traj, perm, weight_path, skpcnf, traj) &
          bind(C, name="erm_traj_open")
       import :: c_ptr, c_int32_t, c_char
       type(c_ptr), value :: ctx, perm, weight_path
       character(kind=c_char), intent(in) :: dcd_path(*)
       integer(c_int32_t), value :: natom_traj, skpcnf
       type(c_ptr), intent(out) :: traj
     end function erm_traj_open_c

     integer(c_int32_t) function erm_traj_next(traj, cell, weight) bind(C, name="erm_traj_next")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: traj
       real(c_double), intent(out) :: cell(3, 3), weight
     end function erm_traj_next

     subroutine erm_traj_close(traj) bind(C, name="erm_traj_close")
       import :: c_ptr
       type(c_ptr), value :: traj
     end subroutine erm_traj_close

     integer(c_int32_t) function erm_read_permutation_c(path, natom_traj, perm, found) &
          bind(C, name="erm_read_permutation")
       import :: c_int32_t, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: natom_traj
       integer(c_int32_t), intent(out) :: perm(*), found
     end function erm_read_permutation_c

     ! several rigid test particles on one solvent as one launch sequence
     ! (no counterpart in the reference: one solute species per run there)
     integer(c_int32_t) function erm_group_create(ctxs, nctx, grp) bind(C, name="erm_group_create")
       import :: c_ptr, c_int32_t
       type(c_ptr), intent(in) :: ctxs(*)
       integer(c_int32_t), value :: nctx
       type(c_ptr), intent(out) :: grp
     end function erm_group_create

     subroutine erm_group_destroy(grp) bind(C, name="erm_group_destroy")
       import :: c_ptr
       type(c_ptr), value :: grp
     end subroutine erm_group_destroy

     type(c_ptr) function erm_group_last_error_c(grp) bind(C, name="erm_group_last_error")
       import :: c_ptr
       type(c_ptr), value :: grp
     end function erm_group_last_error_c

     integer(c_int32_t) function erm_group_set_frame(grp, sitepos, cell, stat_weight_system) &
          bind(C, name="erm_group_set_frame")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: grp
       real(c_double), intent(in) :: sitepos(*), cell(3, 3)
       real(c_double), value :: stat_weight_system
     end function erm_group_set_frame

     integer(c_int32_t) function erm_group_insert_random(grp, n, seed, first_index) &
          bind(C, name="erm_group_insert_random")
       import :: c_ptr, c_int32_t, c_int64_t
       type(c_ptr), value :: grp
       integer(c_int32_t), intent(in) :: n(*)
       integer(c_int64_t), intent(in) :: seed(*), first_index(*)
     end function erm_group_insert_random

     integer(c_int32_t) function erm_group_set_tuning(grp, max_batch, persistent) &
          bind(C, name="erm_group_set_tuning")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: grp
       integer(c_int32_t), value :: max_batch, persistent
     end function erm_group_set_tuning

     integer(c_int32_t) function erm_group_set_graph(grp, on) bind(C, name="erm_group_set_graph")
       import :: c_ptr, c_int32_t
       type(c_ptr), value :: grp
       integer(c_int32_t), value :: on
     end function erm_group_set_graph

     integer(c_int32_t) function erm_group_stage_times(grp, ms10) bind(C, name="erm_group_stage_times")
       import :: c_ptr, c_int32_t, c_double
       type(c_ptr), value :: grp
       real(c_double), intent(out) :: ms10(10)
     end function erm_group_stage_times
  end interface

contains

  ! message of the last failing call as a Fortran string
  function erm_last_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    character(kind=c_char), pointer :: s(:)
    integer :: i
    msg = ''
    if (.not. c_associated(erm_last_error_c(ctx))) return
    call c_f_pointer(erm_last_error_c(ctx), s, [512])
    do i = 1, 512
       if (s(i) == c_null_char) exit
       msg(i:i) = s(i)
    end do
  end function erm_last_error

  ! a DCD trajectory as the frame source of ctx; perm0 (0-based) and the
  ! SysWght path are optional
  integer(c_int32_t) function erm_traj_open(ctx, dcd_path, natom_traj, skpcnf, traj, perm0, weight_path) &
       result(st)
    use iso_c_binding, only: c_loc
    type(c_ptr), intent(in) :: ctx
    character(*), intent(in) :: dcd_path
    integer(c_int32_t), intent(in) :: natom_traj, skpcnf
    type(c_ptr), intent(out) :: traj
    integer(c_int32_t), intent(in), target, optional :: perm0(:)
    character(*), intent(in), optional :: weight_path
    character(kind=c_char, len=:), allocatable, target :: wp
    type(c_ptr) :: pp, wpp
    pp = c_null_ptr
    wpp = c_null_ptr
    if (present(perm0)) pp = c_loc(perm0)
    if (present(weight_path)) then
       wp = trim(weight_path) // c_null_char
       wpp = c_loc(wp)
    end if
    st = erm_traj_open_c(ctx, trim(dcd_path) // c_null_char, natom_traj, pp, wpp, skpcnf, traj)
  end function erm_traj_open

  ! PermIndex into the 0-based perm0; found = 0 (identity) without the file
  integer(c_int32_t) function erm_read_permutation(path, natom_traj, perm0, found) result(st)
    character(*), intent(in) :: path
    integer(c_int32_t), intent(in) :: natom_traj
    integer(c_int32_t), intent(out) :: perm0(*), found
    st = erm_read_permutation_c(trim(path) // c_null_char, natom_traj, perm0, found)
  end function erm_read_permutation

  ! per-frame fluctuation output of engconst (engproc.F90:394-438) on an open
  ! unit: for solution systems (slttype 1) flcuv.tt, the per-species sums of
  ! the solute-solvent energy of every solute, '(i9, 9999f15.5)' with the
  ! frame number, or '(2i9, 9999f15.5)' with the solute's index and the frame
  ! number when there are several; for test particles progress.tt, the frame
  ! number.  uvall(0:slvmax, nconf) as erm_engconst returns it; specofk(k) is
  ! uvspec(tagpt(k)) (0 for the solute itself, whose energy is skipped).
  subroutine erm_write_flcuv(unit, slttype, frame_no, nconf, slvmax, numslv, uvall, specofk)
    integer, intent(in) :: unit, slttype, frame_no, nconf, slvmax, numslv
    real(c_double), intent(in), optional :: uvall(0:slvmax, *)
    integer(c_int32_t), intent(in), optional :: specofk(*)
    real(c_double) :: flc(numslv)
    integer :: cnt, k, pti
    if (slttype /= 1) then
       write(unit, *) frame_no
       return
    end if
    do cnt = 1, nconf
       flc(:) = 0.0d0
       do k = 1, slvmax
          pti = specofk(k)
          if (pti <= 0) cycle
          flc(pti) = flc(pti) + uvall(k, cnt)
       end do
       if (nconf == 1) then
          write(unit, '(i9, 9999f15.5)') frame_no, flc
       else
          write(unit, '(2i9, 9999f15.5)') cnt, frame_no, flc
       end if
    end do
  end subroutine erm_write_flcuv

  ! the bins empty in the reference solvent (extensions/Intermediate/
  ! engproc.F90:286-315): refs_filename.tt, or every existing .01 ... .99
  ! file, each of ermax lines whose third value is the density; a bin is
  ! empty unless some file gives it more than 1e-30.  Returns ERM_E_ARG when
  ! no file exists or one is short.
  integer(c_int32_t) function erm_read_zerodns(refs_filename, ermax, zerodns) result(st)
    character(*), intent(in) :: refs_filename
    integer, intent(in) :: ermax
    integer(c_int32_t), intent(out) :: zerodns(ermax)
    character(len=len_trim(refs_filename) + 3) :: fname
    character(len=3) :: suf
    logical :: ex
    integer :: nfile, k, iduv, u, idum, ios, nread
    real(c_double) :: factor
    st = ERM_OK
    zerodns(:) = 1
    fname = trim(refs_filename) // '.tt'
    inquire(file=fname, exist=ex)
    nfile = 99
    if (ex) nfile = 1
    nread = 0
    do k = 1, nfile
       if (nfile > 1) then
          write(suf, '(a1,i2.2)') '.', k
          fname = trim(refs_filename) // suf
          inquire(file=fname, exist=ex)
       end if
       if (.not. ex) cycle
       nread = nread + 1
       open(newunit=u, file=fname, action='read')
       do iduv = 1, ermax
          read(u, *, iostat=ios) factor, idum, factor
          if (ios /= 0) then
             st = ERM_E_ARG
             close(u)
             return
          end if
          if (factor > 1.0d-30) zerodns(iduv) = 0
       end do
       close(u)
    end do
    if (nread == 0) st = ERM_E_ARG
  end function erm_read_zerodns

  ! the intermediate state's output of engstore (extensions/Intermediate/
  ! engproc.F90:655-694), after erm_engstore: engint.NN ('.tt' for one
  ! division) in the engsln format; at the last division Prob_Result (the
  ! probability soln_intm / soln_total and temp * log of it) and avint.tt.
  ! avintm(engdiv, numslv) carries the per-division averages across calls.
  subroutine erm_store_intermediate(ctx, division, engdiv, temp, numslv, uvmax, uvsoft, uvcrd, avintm, dir)
    use iso_c_binding, only: c_loc
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: division, engdiv, numslv
    integer(c_int32_t), intent(in) :: uvmax(numslv), uvsoft(numslv)
    real(c_double), intent(in) :: temp, uvcrd(*)
    real(c_double), intent(inout) :: avintm(engdiv, numslv)
    character(*), intent(in) :: dir
    real(c_double), allocatable :: edintm(:), uvintm(:)
    real(c_double) :: sc(3), factor
    integer(c_int32_t) :: slvmax, ermax, st
    character(len=3) :: suff
    integer :: u, iduv, idmin, pti, k
    st = erm_sizes(ctx, slvmax, ermax)
    allocate(edintm(ermax), uvintm(0:numslv))
    st = erm_fetch_intermediate(ctx, edintm, uvintm, sc)
    if (engdiv == 1) then
       suff = '.tt'
    else
       write(suff, '(a1,i2.2)') '.', mod(division, 100)
    end if
    open(newunit=u, file=dir//'/engint'//suff, form='formatted', action='write')
    idmin = 0
    pti = 1
    do iduv = 1, ermax
       do while (iduv > idmin + uvmax(pti))
          idmin = idmin + uvmax(pti)
          pti = pti + 1
       end do
       factor = erm_repval(uvcrd(idmin + 1), uvmax(pti), uvsoft(pti), iduv - idmin)
       write(u, '(g15.7,i5,g25.15)') factor, pti, edintm(iduv)
    end do
    endfile(u)
    close(u)
    avintm(division, 1:numslv) = uvintm(1:numslv) / sc(1)
    if (division == engdiv) then
       factor = sc(2) / sc(3)
       open(newunit=u, file=dir//'/Prob_Result', action='write')
       write(u, '(a,f12.6)') ' Probability to find the system in the intermediate state = ', factor
       write(u, '(a)') ' Free-energy change from the intermediate state to the simulated state'
       write(u, '(a,f12.6,a)') '                                                = ', temp * log(factor), &
            ' kcal/mol'
       endfile(u)
       close(u)
       open(newunit=u, file=dir//'/avint.tt', action='write')
       do k = 1, engdiv
          write(u, '(i5, 9999f15.5)') k, avintm(k, 1:numslv)
       end do
       endfile(u)
       close(u)
    end if
  end subroutine erm_store_intermediate

  ! engstore's file output (engproc.F90:560-656), called after erm_engstore
  ! for division `division` of `engdiv`: engsln|engref.NN and, with corrcal,
  ! corsln|corref.NN, with selfcal slfeng.NN ('.tt' when engdiv = 1); at the
  ! last division aveuv.tt (soln), weight_soln|weight_refs and uvrange.tt.
  ! avediv(engdiv, 2) and aveuv(engdiv, numslv) carry the per-division
  ! values across calls.  Files go to directory `dir`.
  subroutine erm_store_division(ctx, division, engdiv, slttype, wgtslf, corrcal, temp, numslv, uvmax, &
       uvsoft, uvcrd, esmax, escrd, avediv, aveuv, dir)
    use iso_c_binding, only: c_loc
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: division, engdiv, slttype, wgtslf, corrcal, numslv, esmax
    integer(c_int32_t), intent(in) :: uvmax(numslv), uvsoft(numslv)
    real(c_double), intent(in) :: temp, uvcrd(*), escrd(*)
    real(c_double), intent(inout) :: avediv(engdiv, 2), aveuv(engdiv, numslv)
    character(*), intent(in) :: dir
    integer(c_int32_t) :: slvmax, ermax, st
    real(c_double), allocatable, target :: edens(:), ecorr(:, :), eself(:), minuv(:), maxuv(:), slnuv(:)
    real(c_double), target :: scal(4)
    real(c_double) :: factor
    character(len=3) :: suff
    character(len=6) :: stem
    integer :: k, iduv, pti, idmin, u
    st = erm_sizes(ctx, slvmax, ermax)
    allocate(edens(ermax), eself(max(esmax, 1)), minuv(0:numslv), maxuv(0:numslv), slnuv(0:numslv))
    if (corrcal /= 0) then
       allocate(ecorr(ermax, ermax))
       st = erm_fetch(ctx, c_loc(edens), c_loc(ecorr), c_loc(eself), c_loc(minuv), c_loc(maxuv), &
                      c_loc(slnuv), c_loc(scal))
    else
       st = erm_fetch(ctx, c_loc(edens), c_null_ptr, c_loc(eself), c_loc(minuv), c_loc(maxuv), &
                      c_loc(slnuv), c_loc(scal))
    end if
    ! scal = engnorm, engsmpl (raw), avslf / engnorm, voffset
    avediv(division, 1) = scal(1) / scal(2)
    if (slttype == 1) then
       aveuv(division, 1:numslv) = slnuv(1:numslv) / scal(1)
       avediv(division, 2) = scal(4) - temp * log(scal(3))
    else
       avediv(division, 2) = scal(4) + temp * log(scal(3))
    end if
    if (division == engdiv) then
       if (slttype == 1) then
          open(newunit=u, file=dir//'/aveuv.tt', action='write')
          do k = 1, engdiv
             write(u, '(i5, 9999f15.5)') k, aveuv(k, 1:numslv)
          end do
          endfile(u)
          close(u)
          open(newunit=u, file=dir//'/weight_soln', action='write')
       else
          open(newunit=u, file=dir//'/weight_refs', action='write')
       end if
       do k = 1, engdiv
          if (wgtslf == 0) then
             write(u, '(i5,e20.8)') k, avediv(k, 1)
          else
             write(u, '(i5,e20.8,e20.7)') k, avediv(k, 1), avediv(k, 2)
          end if
       end do
       endfile(u)
       close(u)
       open(newunit=u, file=dir//'/uvrange.tt', action='write')
       write(u, '(A)') ' species     minimum        maximum'
       do pti = 0, numslv
          if (maxuv(pti) < 1.0d5) then
             write(u, '(i5,2f15.5)') pti, minuv(pti), maxuv(pti)
          else
             write(u, '(i5,f15.5,g18.5)') pti, minuv(pti), maxuv(pti)
          end if
       end do
       endfile(u)
       close(u)
    end if
    if (engdiv == 1) then
       suff = '.tt'
    else
       write(suff, '(a1,i2.2)') '.', mod(division, 100)
    end if
    stem = 'engref'
    if (slttype == 1) stem = 'engsln'
    open(newunit=u, file=dir//'/'//stem//suff, form='formatted', action='write')
    idmin = 0
    pti = 1
    do iduv = 1, ermax
       do while (iduv > idmin + uvmax(pti))
          idmin = idmin + uvmax(pti)
          pti = pti + 1
       end do
       ! the species' edges from uvcrd(idmin + 1) on (sequence association,
       ! like the reference's uvcrd(iduv + 1) at the last linear bin)
       factor = erm_repval(uvcrd(idmin + 1), uvmax(pti), uvsoft(pti), iduv - idmin)
       write(u, '(g15.7,i5,g25.15)') factor, pti, edens(iduv)
    end do
    endfile(u)
    close(u)
    if (corrcal /= 0) then
       stem = 'corref'
       if (slttype == 1) stem = 'corsln'
       open(newunit=u, file=dir//'/'//stem//suff, form='unformatted', action='write')
       write(u) ecorr
       endfile(u)
       close(u)
    end if
    if (esmax > 0) then
       open(newunit=u, file=dir//'/slfeng'//suff, form='formatted', action='write')
       do iduv = 1, esmax
          if (iduv < esmax) then
             factor = (escrd(iduv) + escrd(iduv + 1)) / 2.0d0
          else
             factor = escrd(esmax)
          end if
          write(u, '(g15.7,g25.15)') factor, eself(iduv)
       end do
       endfile(u)
       close(u)
    end if
  end subroutine erm_store_division

  ! representative energy of bin iduv (1-based) of one species, as repval
  ! (engproc.F90:1185-1223) computes it for the engsln/engref files
  real(c_double) function erm_repval(crd, idmax, idsoft, idpt) result(v)
    real(c_double), intent(in) :: crd(*)
    integer, intent(in) :: idmax, idsoft, idpt
    if (idpt <= idsoft) then
       v = (crd(idpt) + crd(idpt + 1)) / 2.0d0
    else if (idpt < idmax) then
       v = sqrt(crd(idpt) * crd(idpt + 1))
    else
       v = crd(idpt)
    end if
  end function erm_repval
end module ermgpu
