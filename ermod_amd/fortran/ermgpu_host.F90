! ermgpu_host.F90 -- Fortran host of the energy pipeline through the
! ISO_C_BINDING module ermgpu: the division loop of ERmod's trjmain
! (enganal.F90:88-97) -- per division engclear, per frame set_frame and
! engconst (random insertions of the test particle, or every solute molecule
! of a solution system) with its flcuv.tt / progress.tt line, engstore and
! the reference's output files (erm_store_division, engproc.F90:560-656).
!
! usage: ermgpu_host <input stream file> <output directory> [trajectory.dcd]
! The input holds, in order and without padding: natom nmol ntype nframes
! nins ndiv (int32); the erm_params fields except pme_path (cltype and
! boxshp last); the erm_hist fields of species 1
! (also used for the self-energy coordinate when selfcal = 1);
! mol_begin(nmol+1) charge(natom) ljtype(natom) lj_eps(ntype^2)
! lj_sig2(ntype^2) sluvid(nmol) uvspec(nmol) mol_charge(nmol)
! sitemass(natom) bfcoord(3*ns) cell(9) seed(int64); then nframes frames
! of 3*natom coordinates, split evenly over the ndiv divisions.  With a DCD
! file the frames (and their cells) come from it instead, through
! erm_traj_*, with <output directory>/PermIndex and SysWght when present.
! For a solution system, a file <output directory>/intermediate holding the
! reference's refs_filename switches the intermediate-state method on.
program ermgpu_host
  use iso_c_binding
  use ermgpu
  implicit none
  type(erm_params) :: prm
  type(erm_hist) :: hst
  type(c_ptr) :: ctx
  integer(c_int32_t) :: natom, nmol, ntype, nframes, nins, ndiv, st, pemax, pesoft, slvmax, ermax, ns
  integer(c_int32_t) :: esmax, essoft
  integer(c_int32_t), allocatable :: mol_begin(:), ljtype(:), sluvid(:), uvspec(:)
  real(c_double), allocatable :: charge(:), lj_eps(:), lj_sig2(:), mol_charge(:), sitemass(:), bfcoord(:)
  real(c_double), allocatable, target :: pos(:), uv(:, :), edens(:), ecorr(:, :), ercrd(:), escrd(:)
  integer(c_int32_t), allocatable, target :: sl0(:)
  integer(c_int32_t), allocatable :: specofk(:)
  integer(c_int32_t) :: nslt, ncfg, k, m, uflc
  real(c_double), allocatable :: avediv(:, :), aveuv(:, :), avintm(:, :)
  integer(c_int32_t), allocatable :: zerodns(:)
  character(len=1024) :: refsname
  logical :: intm
  real(c_double), target :: scal(4)
  real(c_double) :: cell(3, 3)
  integer(c_int64_t) :: seed
  integer :: f, div, fpd
  character(len=1024) :: inp, outdir, dcd
  type(c_ptr) :: traj
  integer(c_int32_t) :: ntraj, found
  integer(c_int32_t), allocatable :: perm0(:)
  real(c_double) :: fw
  logical :: use_dcd, have_w

  call get_command_argument(1, inp)
  call get_command_argument(2, outdir)
  call get_command_argument(3, dcd)
  use_dcd = len_trim(dcd) > 0
  open(10, file=trim(inp), access='stream', form='unformatted', status='old', action='read')
  read(10) natom, nmol, ntype, nframes, nins, ndiv
  read(10) prm%slttype, prm%estype, prm%ljswitch, prm%splodr, prm%ms, prm%corrcal, prm%selfcal, &
       prm%wgtslf, prm%force_calculation, prm%elecut, prm%lwljcut, prm%upljcut, prm%screen, &
       prm%temp, prm%block_threshold, prm%precision, prm%max_batch, prm%device, prm%cltype, prm%boxshp
  prm%pme_path = 0
  read(10) hst%ecdmin, hst%ecfmns, hst%ecdcen, hst%eccore, hst%ecdmax, hst%eclbin, hst%ecfbin, &
       hst%ec0bin, hst%finfac, hst%pecore
  allocate(mol_begin(nmol + 1), charge(natom), ljtype(natom), lj_eps(ntype * ntype), &
       lj_sig2(ntype * ntype), sluvid(nmol), uvspec(nmol), mol_charge(nmol), sitemass(natom))
  read(10) mol_begin, charge, ljtype, lj_eps, lj_sig2, sluvid, uvspec, mol_charge, sitemass
  ns = mol_begin(nmol + 1) - mol_begin(nmol)
  allocate(bfcoord(3 * ns), pos(3 * natom))
  read(10) bfcoord, cell, seed

  ! enginit: bin edges of the single solvent species (and of the self energy)
  allocate(ercrd(10002), escrd(10002))
  st = erm_build_edges(hst, 0, ercrd, 10002, pemax, pesoft)
  call check(st, 'erm_build_edges')
  esmax = 0
  if (prm%selfcal /= 0) then
     st = erm_build_edges(hst, 1, escrd, 10002, esmax, essoft)
     call check(st, 'erm_build_edges (self)')
  end if
  st = erm_create(prm, ctx)
  call check(st, 'erm_create')
  st = erm_set_topology(ctx, natom, nmol, mol_begin, charge, ljtype, ntype, lj_eps, lj_sig2, &
       sluvid, uvspec, mol_charge, sitemass, bfcoord)
  call check(st, 'erm_set_topology')
  st = erm_set_bins(ctx, 1, [pemax], [pesoft], ercrd, esmax, escrd)
  call check(st, 'erm_set_bins')
  st = erm_sizes(ctx, slvmax, ermax)
  ! solutes (sltlist) and the species of every physical molecule (tagpt)
  nslt = count(sluvid == 1)
  allocate(sl0(max(nslt, 1)), specofk(slvmax))
  k = 0
  nslt = 0
  do m = 1, nmol
     if (sluvid(m) > 1) cycle
     k = k + 1
     specofk(k) = uvspec(m)
     if (sluvid(m) == 1) then
        nslt = nslt + 1
        sl0(nslt) = m - 1
     end if
  end do
  ncfg = nins
  if (prm%slttype == 1) ncfg = nslt
  allocate(uv(slvmax + 1, ncfg), edens(ermax), ecorr(ermax, ermax))
  if (prm%slttype == 1) then
     open(newunit=uflc, file=trim(outdir)//'/flcuv.tt', status='replace', action='write')
  else
     open(newunit=uflc, file=trim(outdir)//'/progress.tt', status='replace', action='write')
  end if
  allocate(avediv(ndiv, 2), aveuv(ndiv, 1), avintm(ndiv, 1))
  inquire(file=trim(outdir)//'/intermediate', exist=intm)
  intm = intm .and. prm%slttype == 1
  if (intm) then
     open(12, file=trim(outdir)//'/intermediate', action='read')
     read(12, '(a)') refsname
     close(12)
     allocate(zerodns(ermax))
     st = erm_read_zerodns(trim(refsname), int(ermax), zerodns)
     call check(st, 'erm_read_zerodns')
     st = erm_set_intermediate(ctx, zerodns)
     call check(st, 'erm_set_intermediate')
  end if
  if (use_dcd) then
     ! the trajectory holds every atom but the test particle's
     ntraj = natom - ns
     allocate(perm0(ntraj))
     st = erm_read_permutation(trim(outdir)//'/PermIndex', ntraj, perm0, found)
     call check(st, 'erm_read_permutation')
     inquire(file=trim(outdir)//'/SysWght', exist=have_w)
     if (found /= 0 .and. have_w) then
        st = erm_traj_open(ctx, trim(dcd), ntraj, 1, traj, perm0, trim(outdir)//'/SysWght')
     else if (found /= 0) then
        st = erm_traj_open(ctx, trim(dcd), ntraj, 1, traj, perm0)
     else if (have_w) then
        st = erm_traj_open(ctx, trim(dcd), ntraj, 1, traj, weight_path=trim(outdir)//'/SysWght')
     else
        st = erm_traj_open(ctx, trim(dcd), ntraj, 1, traj)
     end if
     call check(st, 'erm_traj_open')
  end if

  fpd = nframes / ndiv
  do div = 1, ndiv
     st = erm_engclear(ctx)
     call check(st, 'erm_engclear')
     do f = (div - 1) * fpd + 1, div * fpd
        if (use_dcd) then
           st = erm_traj_next(traj, cell, fw)
           call check(st, 'erm_traj_next')
        else
           read(10) pos
           st = erm_set_frame(ctx, pos, cell, 1.0d0)
           call check(st, 'erm_set_frame')
        end if
        if (prm%slttype == 1) then
           st = erm_engconst(ctx, nslt, c_loc(sl0), c_null_ptr, c_null_ptr, c_loc(uv))
           call check(st, 'erm_engconst')
           call erm_write_flcuv(uflc, 1, f, int(nslt), int(slvmax), 1, uv, specofk)
        else
           if (f == 1) then
              st = erm_insert_random(ctx, nins, seed, int((f - 1) * nins, c_int64_t), c_loc(uv))
           else
              st = erm_insert_random(ctx, nins, seed, int((f - 1) * nins, c_int64_t), c_null_ptr)
           end if
           call check(st, 'erm_insert_random')
           call erm_write_flcuv(uflc, int(prm%slttype), f, 0, int(slvmax), 1)
        end if
     end do
     if (div == ndiv) then
        ! raw accumulators of the last division, before normalisation
        ecorr = 0.0d0
        if (prm%corrcal /= 0) then
           st = erm_fetch(ctx, c_loc(edens), c_loc(ecorr), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                          c_loc(scal))
        else
           st = erm_fetch(ctx, c_loc(edens), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                          c_loc(scal))
        end if
        call check(st, 'erm_fetch')
        open(11, file=trim(outdir)//'/raw.bin', access='stream', form='unformatted', status='replace')
        write(11) slvmax, ermax, scal, edens, ecorr, uv
        close(11)
     end if
     st = erm_engstore(ctx)
     call check(st, 'erm_engstore')
     call erm_store_division(ctx, div, int(ndiv), int(prm%slttype), int(prm%wgtslf), int(prm%corrcal), &
          prm%temp, 1, [pemax], [pesoft], ercrd, int(esmax), escrd, avediv, aveuv, trim(outdir))
     if (intm) call erm_store_intermediate(ctx, div, int(ndiv), prm%temp, 1, [pemax], [pesoft], ercrd, &
          avintm, trim(outdir))
  end do
  close(10)
  endfile(uflc)
  close(uflc)
  if (use_dcd) call erm_traj_close(traj)
  call erm_destroy(ctx)
  print '(a,i0,a,i0,a,i0,a)', 'ermgpu_host: ', ndiv, ' divisions, ', nframes, ' frames x ', nins, &
       ' insertions done'

contains
  subroutine check(code, what)
    integer(c_int32_t), intent(in) :: code
    character(*), intent(in) :: what
    if (code /= ERM_OK) then
       print '(a,a,a,i0,a,a)', 'ermgpu_host: ', what, ' failed with ', code, ': ', trim(erm_last_error(ctx))
       stop 1
    end if
  end subroutine check
end program ermgpu_host
