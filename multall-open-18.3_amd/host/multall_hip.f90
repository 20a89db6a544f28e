!> Driver program: the reference's main program (multall-open-18.3.f:89-142)
!> with `CALL LOOP` replaced by the MI355X library.  Reads a NEW-format deck
!> on standard input, builds the grid with mh_host, uploads it, and marches
!> in chunks of 5 steps -- writing the reference's convergence line to
!> stage.log (mf:5437-5438) and polling the `stopit` file between chunks; a
!> request ends the run at the next step divisible by 10, the step at which
!> the reference reads the file (mf:3774-3779) -- until NSTEPS_MAX or
!> convergence.  Writes `flow_out` in
!> the reference's record layout (mf:5951-5976) at the end.
program multall_hip
  use iso_c_binding
  use mh_host
  use mh_hipif
  implicit none
  type(mh_dims) :: dims
  type(mh_resid) :: res
  type(c_ptr) :: ctx
  integer :: rc, ifstop, ios, nstep_done
  character(len=1) :: ansin
  real, allocatable, target :: buf(:,:,:)

  open (unit=13, file='intype', status='old', iostat=ios)
  if (ios == 0) then
    read (13, *, iostat=ios) ansin
    close (13)
    if (ansin == 'O' .or. ansin == 'o') then
      write (*, *) ' OLD_READIN decks are not supported; use the NEW format.'
      stop 2
    end if
  end if
  open (unit=12, file='stopit')
  write (12, *) 0
  close (12)
  ! the wall distances and mixing lengths are formed on the device below
  host_xlength = .false.
  if (mh_host_load_unit(5)) then
    write (*, *) ' DECK/SETUP FAILED: ', trim(mh_host_err)
    stop 2
  end if
  write (*, '(a,3i6,a,i3)') ' GRID IM,JM,KM =', IM, JM, KM, '   ROWS =', NROWS
  ! the plotting grid file SETUP writes, mf:6585-6599
  if (write_grid_out('grid_out')) write (*, *) ' COULD NOT WRITE grid_out: ', trim(mh_host_err)
  dims = mh_dims(IM, JM, KM, NROWS, NIB1, NJB1, NKB1, NIB2, NJB2, NKB2, NSBLK)
  rc = mh_create(dims, ctx)
  call check(rc, 'mh_create')
  call upload()
  ! SETUP's SET_XLENGTH (mf:7340) on the device, then its initial viscous
  ! forces (mf:7346-7357); both no-ops when ILOS = 0
  if (ILOS /= 0) then
    rc = mh_set_xlength(ctx, THETA, XLLIM_IN, XLLIM_LE, XLLIM_TE, XLLIM_DN)
    call check(rc, 'mh_set_xlength')
  end if
  rc = mh_loss_init(ctx)
  call check(rc, 'mh_loss_init')
  rc = mh_start(ctx)
  call check(rc, 'mh_start')
  open (unit=4, file='stage.log')
  write (*, *) ' STARTING THE MAIN TIME STEPPING LOOP ON THE GPU'
  nstep_done = 0
  do
    rc = mh_step(ctx, 5, res)
    call check(rc, 'mh_step')
    if (res%nstep == nstep_done) exit
    nstep_done = res%nstep
    ! the convergence line of every check, mf:5437-5438
    write (4, '(a,3e12.5,a,e12.5,a,i5,a,3i5)') 'EMAX,EAVG,ECONT', res%emax, res%eavg, res%econt, ' FLOW', &
      res%flow_in, 'STEP', res%nstep, 'AT', res%imax, res%jmax, res%kmax
    if (mod(res%nstep, 50) == 0 .or. res%ifend == 1) &
      write (*, '(i6,3f9.4,2f10.2,2f12.4)') res%nstep, res%emax, res%eavg, res%econt, res%vref, res%vmax, &
      res%flow_in, res%flow_out
    if (res%ifend == 1) exit
    open (unit=12, file='stopit')
    read (12, *, iostat=ios) ifstop
    close (12)
    if (ios == 0 .and. ifstop >= 1) then
      rc = mh_request_stop(ctx)
      call check(rc, 'mh_request_stop')
    end if
  end do
  close (4)
  call write_flow_out()
  call print_performance()
  call write_mixbconds()
  rc = mh_destroy(ctx)
  write (*, *) ' STOPPED AFTER ', res%nstep, ' STEPS'

contains

  subroutine check(r, what)
    integer, intent(in) :: r
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (r == 0) return
    call c_f_pointer(mh_last_error(ctx), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write (*, *) ' LIBRARY ERROR in ', what, ' rc=', r, ': ', msg(1:n)
    stop 3
  end subroutine check

  subroutine put3(name, a)
    character(len=*), intent(in) :: name
    real, intent(in), target, contiguous :: a(:,:,:)
    call check(mh_set_array(ctx, mh_array_id(cname(name)), c_loc(a), int(size(a), c_long_long)), name)
  end subroutine put3
  subroutine put2(name, a)
    character(len=*), intent(in) :: name
    real, intent(in), target, contiguous :: a(:,:)
    call check(mh_set_array(ctx, mh_array_id(cname(name)), c_loc(a), int(size(a), c_long_long)), name)
  end subroutine put2
  subroutine put1(name, a)
    character(len=*), intent(in) :: name
    real, intent(in), target, contiguous :: a(:)
    call check(mh_set_array(ctx, mh_array_id(cname(name)), c_loc(a), int(size(a), c_long_long)), name)
  end subroutine put1
  subroutine puti(name, a)
    character(len=*), intent(in) :: name
    integer, intent(in), target, contiguous :: a(:)
    call check(mh_set_array(ctx, mh_array_id(cname(name)), c_loc(a), int(size(a), c_long_long)), name)
  end subroutine puti
  subroutine sets(name, v)
    character(len=*), intent(in) :: name
    real(c_double), intent(in) :: v
    call check(mh_set_scalar(ctx, mh_scalar_id(cname(name)), v), name)
  end subroutine sets

  subroutine upload()
    call put3('RO', RO); call put3('ROVX', ROVX); call put3('ROVR', ROVR); call put3('RORVT', RORVT)
    call put3('ROE', ROE); call put3('DRO', DRO); call put3('DROVX', DROVX); call put3('DROVR', DROVR)
    call put3('DRORVT', DRORVT); call put3('DROE', DROE); call put3('ROVT', ROVT); call put3('ROWT', ROWT)
    call put3('P', P); call put3('HO', HO); call put3('PEFF', PEFF); call put3('VX', VX); call put3('VR', VR)
    call put3('VT', VT); call put3('WT', WT); call put3('T_STATIC', T_STATIC); call put3('AQX', AQX)
    call put3('AQR', AQR); call put3('ABX', ABX); call put3('ABR', ABR); call put3('ASX', ASX)
    call put3('ASR', ASR); call put3('VOL', VOL); call put3('VOLOR', VOLOR); call put3('STEP', STEP)
    call put3('BSTEP', BSTEP); call put3('RSTEP', RSTEP); call put3('XFORCE', XFORCE)
    call put3('RFORCE', RFORCE); call put3('TFORCE', TFORCE); call put3('QSOURCE', QSOURCE)
    call put3('FLOWX', FLOWX); call put3('FLOWT', FLOWT); call put3('FLOWR', FLOWR)
    call put2('R', R); call put2('X', X); call put2('RAVG_CELL', RAVG_CELL); call put2('ABT', ABT)
    call put2('WRABT', WRABT); call put2('UBLADE', UBLADE); call put2('DX', DX); call put2('DR', DR)
    call put2('DS', DS); call put2('SMERID', SMERID); call put2('FTL', FTL); call put2('FBL', FBL)
    call put2('FTR', FTR); call put2('FBR', FBR)
    call put1('FP', FP(1:IM)); call put1('FU', FU(1:IM)); call put1('FD', FD(1:IM))
    call put1('FR', FR(1:KM)); call put1('FKU', FKU(1:KM)); call put1('FKD', FKD(1:KM))
    call put1('PO1', PO1(1:KM)); call put1('TO1', TO1(1:KM)); call put1('RPO1', RPO1(1:KM))
    call put1('BSSIN', BSSIN(1:KM)); call put1('BSCOS', BSCOS(1:KM)); call put1('BRSIN', BRSIN(1:KM))
    call put1('BRCOS', BRCOS(1:KM)); call put1('VTIN', VTIN(1:KM)); call put1('PD', PD(1:KM))
    call put1('VM1', VM1(1:KM)); call put1('FACUP', FACUP); call put1('FACDWN', FACDWN)
    call put1('WRAD', WRAD); call put3('STEP1', STEP1); call put3('STEP2', STEP2); call put1('STEPSBK', STEPSBK)
    call puti('IB1', IB1); call puti('IB2', IB2); call puti('KB1', KB1); call puti('KB2', KB2)
    call puti('NBLADE', NBLADE); call puti('IND', IND); call puti('INDLE', INDLE); call puti('INDMIX', INDMIX)
    call puti('INDTE', INDTE); call puti('NROW', NROW); call puti('JSBLK', JSBLK); call puti('JB1', JB1)
    call puti('JB2', JB2); call puti('JSTART', JSTART(1:NROWS + 1)); call puti('JMIX', JMIX(1:NROWS + 1))
    call puti('JLE', JLE(1:NROWS + 1)); call puti('JTE', JTE(1:NROWS + 1)); call puti('JLED', JLED(1:NROWS + 1))
    call puti('KTIPS', KTIPS(1:NROWS + 1)); call puti('KTIPE', KTIPE(1:NROWS + 1))
    call puti('KS1', KS1(1:NROWS + 1)); call puti('KS2', KS2(1:NROWS + 1))
    call puti('NOUT', NOUT); call puti('NSTAGE', NSTAGE(1:NROWS))
    ! the trailing-edge separation force of IF_CUSP = 2 rows
    call puti('IF_CUSP', IF_CUSP(1:NROWS)); call puti('NUP_I1', NUP_I1); call puti('NUP_IM', NUP_IM)
    call puti('N_WAKE', N_WAKE); call put1('SEP_THIK', SEP_THIK); call put1('SEP_DRAG', SEP_DRAG)
    call put2('RT_THICK', RT_THICK(1:JM, 1:KM)); call put3('RTHETA', RTHETA)
    call sets('CFL', dble(CFL)); call sets('DAMPIN', dble(DAMPIN)); call sets('MACHLIM', dble(MACHLIM))
    call sets('F_PDOWN', dble(F_PDOWN)); call sets('CONLIM', dble(CONLIM)); call sets('SFXIN', dble(SFXIN))
    call sets('SFTIN', dble(SFTIN)); call sets('FAC_4TH', dble(FAC_4TH)); call sets('F1', dble(F1))
    call sets('F2', dble(F2)); call sets('F3', dble(F3)); call sets('RSMTH', dble(RSMTH))
    call sets('SFEXIT', dble(SFEXIT)); call sets('FSMTHB', dble(FSMTHB)); call sets('RFIN', dble(RFIN))
    call sets('FLOWIN', dble(FLOWIN)); call sets('RFLOW', dble(RFLOW))
    ! the repeating-stage condition
    call sets('NINMOD', dble(NINMOD)); call sets('RFINBC', dble(RFINBC))
    call sets('PO_IN_MID', dble(PO_IN_MID)); call sets('VT_IN_MID', dble(VT_IN_MID))
    call put1('BS', BS(1:KM)); call put1('BR', BR(1:KM))
    call sets('PDOWN_HUB', dble(PDOWN_HUB)); call sets('PDOWN_TIP', dble(PDOWN_TIP))
    call sets('FEXTRAP', dble(FEXTRAP)); call sets('RFMIX', dble(RFMIX)); call sets('FANGLE', dble(FANGLE))
    call sets('CP', dble(CP)); call sets('GA', dble(GA)); call sets('RGAS', dble(RGAS)); call sets('CV', dble(CV))
    call sets('FGA', dble(FGA)); call sets('GA1', dble(GA1)); call sets('RCP', dble(RCP))
    call sets('VSOUND', dble(VSOUND)); call sets('FACMIXUP', dble(FACMIXUP)); call sets('IMID', dble(IMID))
    call sets('KMID', dble(KMID)); call sets('NSTEPS_MAX', dble(NSTEPS_MAX)); call sets('NCHANGE', dble(NCHANGE))
    call sets('NRSMTH', dble(NRSMTH)); call sets('NSFEXIT', dble(NSFEXIT)); call sets('IN_PRESS', dble(IN_PRESS))
    call sets('IN_VTAN', dble(IN_VTAN)); call sets('IN_VR', dble(IN_VR)); call sets('IPOUT', dble(IPOUT))
    call sets('IFMIX', dble(IFMIX)); call sets('IR', dble(IR)); call sets('JR', dble(JR)); call sets('KR', dble(KR))
    call sets('ILOS', dble(ILOS)); call sets('NLOS', dble(NLOS)); call sets('NMIXUP', dble(NMIXUP))
    call sets('ITIMST', dble(ITIMST)); call sets('IF_RESTART', dble(IF_RESTART)); call sets('IFGAS', dble(IFGAS))
    call sets('IN_FLOW', dble(IN_FLOW)); call sets('IF_REPEAT', dble(IF_REPEAT)); call sets('IRBB', dble(IRBB))
    call sets('KRBB', dble(KRBB))
    ! viscous model inputs (LOSS)
    call put3('DWALLSQ', DWALLSQ); call put1('WHUB', WHUB); call put1('WTIP', WTIP)
    call put1('XLLIM_I1', XLLIM_I1); call put1('XLLIM_IM', XLLIM_IM); call put1('XLLIM_K1', XLLIM_K1)
    call put1('XLLIM_KM', XLLIM_KM); call put1('XLLIM_DWN', XLLIM_DWN); call put1('XLLIM_UP', XLLIM_UP)
    call put1('ROUGH_H', ROUGH_H); call put1('ROUGH_T', ROUGH_T); call put1('ROUGH_L', ROUGH_L)
    call put1('ROUGH_U', ROUGH_U); call puti('JTRAN_I1', JTRAN_I1); call puti('JTRAN_IM', JTRAN_IM)
    call puti('JTRAN_K1', JTRAN_K1); call puti('JTRAN_KM', JTRAN_KM)
    call put1('CHORD', CHORD)
    call sets('REYNO', dble(REYNO)); call sets('RF_VIS', dble(RF_VIS)); call sets('FTRANS', dble(FTRANS))
    call sets('PRANDTL', dble(PRANDTL)); call sets('YPLUSWALL', dble(YPLUSWALL)); call sets('YPLAM', dble(YPLAM))
    call sets('YPTURB', dble(YPTURB)); call sets('CFWALL', dble(CFWALL)); call sets('IBOUND', dble(IBOUND))
    ! Spalart-Allmaras inputs and the turbulence state SETUP starts from
    ! (mf:6903-6909)
    call put3('XLENGTH', XLENGTH); call put3('FST_RAT', FST_RAT); call put3('ROAVG_CELL', ROAVG_CELL)
    call put3('TRANS_DYN_VIS', TRANS_DYN_VIS)
    call put1('FSTURB', FSTURB); call put1('TURBVIS_DAMP', TURBVIS_DAMP)
    allocate (buf(IM, JM, KM))
    buf = 25.0
    call put3('Y_PLUS', buf)
    buf = 1.0
    call put3('VISC_RAT', buf)
    deallocate (buf)
    call sets('FAC_STMIX', dble(FAC_STMIX)); call sets('FAC_ST0', dble(FAC_ST0)); call sets('FAC_ST1', dble(FAC_ST1))
    call sets('FAC_ST2', dble(FAC_ST2)); call sets('FAC_ST3', dble(FAC_ST3)); call sets('FAC_SFVIS', dble(FAC_SFVIS))
    call sets('FAC_VORT', dble(FAC_VORT)); call sets('FAC_PGRAD', dble(FAC_PGRAD))
    call sets('VORT_MAX', dble(VORT_MAX)); call sets('TURBVIS_LIM', dble(TURBVIS_LIM))
  end subroutine upload

  subroutine get3(name, a)
    character(len=*), intent(in) :: name
    real, intent(inout), target, contiguous :: a(:,:,:)
    call check(mh_get_array(ctx, mh_array_id(cname(name)), c_loc(a), int(size(a), c_long_long)), name)
  end subroutine get3

  ! EFICOOL's stage and machine figures at the end of the run, as the
  ! reference prints them (mf:11497-11548; FP64 on the device)
  subroutine print_performance()
    type(mh_perf) :: pf(NROWS + 1)
    integer(c_int) :: nrec
    integer :: m
    call check(mh_eficool(ctx, pf, int(NROWS + 1, c_int), nrec), 'mh_eficool')
    do m = 1, nrec
      write (*, *)
      if (pf(m)%stage > 0) then
        write (*, *) ' RESULTS FOR STAGE NUMBER ', pf(m)%stage
      else
        write (*, *) ' THE FOLLOWING RESULTS ARE FOR THE WHOLE MACHINE'
      end if
      if (pf(m)%compressor == 0) then
        write (*, *) ' ********** THIS IS A TURBINE ********** '
        write (*, *) ' PRESSURE RATIO, POinlet/POexit            = ', 1.0d0/pf(m)%po_ratio
      else
        write (*, *) ' ***********THIS IS A COMPRESSOR ********** '
        write (*, *) ' PRESSURE RATIO, POexit/POinlet            = ', pf(m)%po_ratio
      end if
      write (*, *) ' EFFICIENCIES BASED ON VALUES AT J = ', pf(m)%j1, pf(m)%j2m1
      write (*, *) ' JMIX  = ', pf(m)%jmix
      write (*, *) ' INLET AND EXIT STAGNATION PRESSURES   = ', pf(m)%po_in, pf(m)%po_out
      write (*, *) ' INLET AND EXIT STATIC PRESSURES       = ', pf(m)%ps_in, pf(m)%ps_out
      write (*, *) ' INLET AND EXIT STAGNATION TEMPERATURES= ', pf(m)%to_in, pf(m)%to_out
      write (*, *) ' INLET AND OUTLET MASS FLOW RATES      = ', pf(m)%flow_in, pf(m)%flow_out
      write (*, *) ' TOTAL TO TOTAL ISENTROPIC EFFICIENCY       = ', pf(m)%eta_tt
      write (*, *) ' TOTAL TO STATIC ISENTROPIC EFFICIENCY      = ', pf(m)%eta_ts
      write (*, *) ' TOTAL TO TOTAL POLYTROPIC EFFICIENCY       = ', pf(m)%eta_poly
      if (pf(m)%stage > 0) then
        write (*, *) ' AVERAGE STAGE REACTION BASED ON MASS AVERAGED TEMPERATURES = ', pf(m)%reaction
        write (*, *) ' HUB REACTION BASED ON PRESSURE CHANGES     ', pf(m)%reaction_hub
        write (*, *) ' MID SPAN REACTION BASED ON PRESSURE CHANGES', pf(m)%reaction_mid
        write (*, *) ' TIP REACTION BASED ON PRESSURE CHANGES     ', pf(m)%reaction_tip
        write (*, *) ' STAGE WORK NEGLECTING ANY BLEED OR COOLING FLOWS = ', pf(m)%work_kw, ' KILOWATTS.'
      else
        write (*, *) ' OVERALL POWER OUTPUT             = ', pf(m)%work_kw, ' KILOWATTS.'
      end if
    end do
  end subroutine print_performance

  ! 'mixbconds' at the end of the run (MIX_BCONDS(1), mf:5445 and
  ! 11906-12080): the mass-averaged spanwise profiles at the inlet and at
  ! each row's JMIX, in the reference's layout; the averages from the device
  subroutine write_mixbconds()
    real(c_double), allocatable :: mb(:)
    integer :: u, nr, q, k, j, base
    real :: fspan(KM)
    character(len=32), parameter :: names(8) = [character(len=32) :: ' STAGNATION PRESSURE ', &
      ' STATIC PRESSURE ', ' STAGNATION TEMPERATURE ', ' STATIC TEMPERATURE ', &
      ' ABSOLUTE TANGENTIAL VELOCITY ', ' MERIDIONAL VELOCITY ', ' YAW ANGLE ', ' PITCH ANGLE ']
    allocate (mb((NROWS + 1)*8*(KM + 1)))
    call check(mh_mix_bconds(ctx, mb, int(size(mb), c_long_long)), 'mh_mix_bconds')
    fspan(1) = 0.0
    do k = 2, KM
      fspan(k) = fspan(k - 1) + FR(k - 1)
    end do
    open (newunit=u, file='mixbconds')
    do nr = 1, NROWS + 1
      write (u, *)
      write (u, *) ' BLADE ROW NUMBER ', nr, ' JMIX = ', JMIX(nr)
      write (u, *)
      if (nr == 1) then
        write (u, *)
        write (u, *) ' INLET BOUNDARY CONDITIONS TO ROW NUMBER 1.'
        write (u, *)
      else
        j = JMIX(nr - 1)
        write (u, *) ' ROW NUMBER ', nr - 1, ' JMIX  =  ', j
        write (u, *) ' EXIT FLOW CONDITIONS FROM THIS BLADE ROW'
        write (u, *) ' THESE CAN BE USED AS THE INLET BOUNDARY CONDITIONS FOR THE NEXT STAGE.'
        write (u, *)
      end if
      do q = 1, 8
        base = ((nr - 1)*8 + q - 1)*(KM + 1)
        write (u, *) trim(names(q))
        if (q <= 4) then
          write (u, '(8F10.1)') (real(mb(base + k)), k = 1, KM)
        else
          write (u, '(8F10.4)') (real(mb(base + k)), k = 1, KM)
        end if
      end do
      write (u, *) ' SPANWISE GRID SPACING '
      write (u, '(8F10.6)') (FR(k), k = 1, KM - 1)
      write (u, *) ' PITCHWISE GRID SPACING '
      write (u, '(8F10.6)') (FP(k), k = 1, IM - 1)
      write (u, *) ' SPANWISE GRID POSITIONS'
      write (u, '(8F10.6)') (fspan(k), k = 1, KM)
    end do
    close (u)
  end subroutine write_mixbconds

  ! flow_out, mf:5951-5976 (ITIMST<5: ROE as marched, ROSUB unset = 0; the
  ! SA variable with ILOS >= 200, else the wall pressure gradient x 1e-6)
  subroutine write_flow_out()
    integer :: u
    call get3('RO', RO); call get3('ROVX', ROVX); call get3('ROVR', ROVR); call get3('ROVT', ROVT)
    call get3('ROE', ROE); call get3('PEFF', PEFF)
    open (newunit=u, file='flow_out', form='unformatted')
    write (u) res%nstep
    write (u) RO
    write (u) ROVX
    write (u) ROVR
    write (u) ROVT
    write (u) ROE
    allocate (buf(IM, JM, KM))
    buf = 0.0
    write (u) buf
    if (ILOS >= 200) then
      call get3('TRANS_DYN_VIS', buf)
    else
      call get3('DPDS_CELL', buf)
      buf = buf*1.0e-6
    end if
    write (u) buf
    write (u) PEFF + PREFF
    close (u)
    deallocate (buf)
  end subroutine write_flow_out

end program multall_hip
