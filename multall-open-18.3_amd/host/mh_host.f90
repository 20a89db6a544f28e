!> Host side of the MI355X time-marching library: reads a Multall NEW-format
!> input deck, builds the H-grid, face areas, time steps, multigrid maps and
!> the initial flow guess, and hands them to libmultall_hip.
!>
!> The statements restate the reference's input and set-up routines
!> (Multall-open-18.3/multall-18.3/multall-open-18.3.f: NEW_READIN :148-2158,
!> INPINT :9073-9195, SETUP :6280-7487, NEW_INTPOL :7493-7610, INTP
!> :7801-7836, STEPUP :7916-8002) in the reference's FP32 statement order,
!> so the arrays handed to the GPU equal the reference's own COMMON state
!> after SETUP bit-for-bit.  Supported deck options: ideal gas, ITIMST 3/-3/4,
!> no cooling, bleed, shroud, restart, throughflow or Q3D, ISHIFT 0 to 4
!> (the inter-row gap grid of GRID_UP/GRID_DOWN, cusps IF_CUSP 0 or 1),
!> INSURF=0, NEWGRID=0, IF_DESIGN/IF_STAGGER/IF_LEAN=0, IF_REPEAT=0, no
!> exit throttle.  Anything else is
!> reported through mh_host_error and the load fails.
!>
!> Arrays are allocated to the logical extents (IM,JM,KM), column-major with
!> I fastest: exactly the dense layout the C ABI takes.
module mh_host
  use iso_c_binding
  implicit none
  private
  public :: mh_host_load_unit, mh_host_load_path, write_grid_out

  character(len=256), save, public :: mh_host_err = ''
  real, parameter :: PI = 3.14159265, DEGRAD = PI/180., RADDEG = 180./PI

  ! ---- scalars (reference names) ------------------------------------------
  integer, save, public :: IM, JM, KM, IMM1, JMM1, KMM1, IMM2, JMM2, KMM2, IMID, KMID, NROWS, NRWSM1
  integer, save, public :: ITIMST, NRSMTH, IF_RESTART, NSTEPS_MAX, NCHANGE, IR, JR, KR, IRBB, JRBB, KRBB
  integer, save, public :: IFMIX, IFCOOL, IFBLEED, IF_ROUGH, NSECS_IN, IN_PRESS, IN_VTAN, IN_VR, IN_FLOW
  integer, save, public :: IF_REPEAT, IPOUT, NSFEXIT, ILOS, NLOS, IBOUND, ISHIFT, NEXTRAP_LE, NEXTRAP_TE
  integer, save, public :: KIN, NMIXUP, NIB1, NJB1, NKB1, NIB2, NJB2, NKB2, NSBLK, IFGAS
  real, save, public :: CP, GA, F1, F2, F3, F2EFF, RSMTH, CFL, DAMPIN, MACHLIM, F_PDOWN, CONLIM
  real, save, public :: SFXIN, SFTIN, FAC_4TH, FBLK1, FBLK2, FBLK3, RFMIX, FEXTRAP, FSMTHB, FANGLE
  real, save, public :: RFIN, SFEXIT, PLATE_LOSS, THROTTLE_EXIT, REYNO, RF_VIS, FTRANS, TURBVIS_LIM
  real, save, public :: FLOWIN = 0.0, RFLOW = 0.0
  ! the repeating-stage condition (IF_REPEAT, mf:575-585): the steps between
  ! its updates and their relaxation factor; the mid-span inlet values the
  ! reader keeps (mf:1889, 1911)
  integer, save, public :: NINMOD = 0
  real, save, public :: RFINBC = 0.0, PO_IN_MID = 0.0, VT_IN_MID = 0.0
  real, save, public :: PRANDTL, YPLUSWALL, YPLAM, YPTURB, PDOWN_HUB, PDOWN_TIP, FACMIXUP, CFWALL
  real, save, public :: FAC_STMIX, FAC_ST0, FAC_ST1, FAC_ST2, FAC_ST3, FAC_SFVIS, FAC_VORT, FAC_PGRAD
  real, save, public :: GA1, FGA, RFGA, RCP, CV, RCV, RGAS, SFEX1, SFEXH, VSOUND, DSMIN, VORT_MAX, PREFF
  integer, save, public :: NOUT(5)
  ! form the wall distances and mixing lengths here (SET_XLENGTH); off when
  ! the caller uses the library's mh_set_xlength instead
  logical, save, public :: host_xlength = .true.

  ! ---- per row --------------------------------------------------------------
  integer, allocatable, save, public :: JSTART(:), JMIX(:), JLE(:), JTE(:), JLED(:), KTIPS(:), KTIPE(:)
  integer, allocatable, save, public :: KS1(:), KS2(:), NSTAGE(:), IF_CUSP(:), IF_ANGLES(:)
  integer, allocatable, save, public :: JTRAN_I1(:), JTRAN_IM(:), JTRAN_K1(:), JTRAN_KM(:)
  real, allocatable, save, public :: FRACTIP1(:), FRACTIP2(:), FTHICK(:,:), CHORD(:)
  integer, allocatable, save, public :: ICUSP(:), LCUSP(:), LCUSPUP(:)
  ! the trailing-edge separation force of IF_CUSP = 2 (mf:993-1004)
  integer, allocatable, save, public :: NUP_I1(:), NUP_IM(:), N_WAKE(:)
  real, allocatable, save, public :: SEP_THIK(:), SEP_DRAG(:)
  real, allocatable, save, public :: BETAUP(:,:), BETADWN1(:,:), BETADWN2(:,:)
  real, allocatable, save, public :: XLLIM_I1(:), XLLIM_IM(:), XLLIM_K1(:), XLLIM_KM(:), XLLIM_DWN(:)
  real, allocatable, save, public :: XLLIM_UP(:), XLLIM_IN(:), XLLIM_LE(:), XLLIM_TE(:), XLLIM_DN(:)
  real, allocatable, save, public :: FSTURB(:), TURBVIS_DAMP(:), ROUGH_H(:), ROUGH_T(:), ROUGH_L(:), ROUGH_U(:)

  ! ---- per J ---------------------------------------------------------------
  integer, allocatable, save, public :: IND(:), INDLE(:), INDTE(:), INDMIX(:), INDMID(:), NBLADE(:), NROW(:)
  integer, allocatable, save, public :: JB1(:), JB2(:), JSBLK(:)
  real, allocatable, save, public :: WRAD(:), WHUB(:), WTIP(:), PGUESS(:), FACUP(:), FACDWN(:)

  ! ---- spanwise / pitchwise vectors -----------------------------------------
  real, allocatable, save, public :: FP(:), FR(:), FU(:), FD(:), FKU(:), FKD(:), FSPAN(:), FR_IN(:)
  real, allocatable, save, public :: PO1(:), TO1(:), RPO1(:), VTIN(:), VM1(:), BS(:), BR(:), PD(:)
  real, allocatable, save, public :: BSSIN(:), BSCOS(:), BRSIN(:), BRCOS(:)
  integer, allocatable, save, public :: IB1(:), IB2(:), KB1(:), KB2(:)

  ! ---- 2D ------------------------------------------------------------------
  real, allocatable, save, public :: XSURF(:,:), RSURF(:,:), X(:,:), R(:,:), RT_UPP(:,:), RT_THICK(:,:)
  real, allocatable, save, public :: SMERID(:,:), DX(:,:), DR(:,:), DS(:,:), RAVG_CELL(:,:), UBLADE(:,:)
  real, allocatable, save, public :: RT_PITCH(:,:), FMUP(:,:), FMDN(:,:), ABT(:,:), WRABT(:,:)
  real, allocatable, save, public :: FTL(:,:), FBL(:,:), FTR(:,:), FBR(:,:)

  ! ---- 3D ------------------------------------------------------------------
  real, allocatable, save, public :: THETA(:,:,:), RTHETA(:,:,:), AQX(:,:,:), AQR(:,:,:), AQTOT(:,:,:)
  real, allocatable, save, public :: ABX(:,:,:), ABR(:,:,:), ABTOT(:,:,:), ASX(:,:,:), ASR(:,:,:)
  real, allocatable, save, public :: VOL(:,:,:), VOLOR(:,:,:), STEP(:,:,:), BSTEP(:,:,:), RSTEP(:,:,:)
  real, allocatable, save, public :: RO(:,:,:), ROVX(:,:,:), ROVR(:,:,:), RORVT(:,:,:), ROE(:,:,:)
  real, allocatable, save, public :: ROVT(:,:,:), ROWT(:,:,:), P(:,:,:), HO(:,:,:), PEFF(:,:,:), VX(:,:,:)
  real, allocatable, save, public :: VR(:,:,:), VT(:,:,:), WT(:,:,:), T_STATIC(:,:,:), TEMP1(:,:,:)
  real, allocatable, save, public :: TEMP2(:,:,:), DRO(:,:,:), DROVX(:,:,:), DROVR(:,:,:), DRORVT(:,:,:)
  real, allocatable, save, public :: DROE(:,:,:), XFORCE(:,:,:), RFORCE(:,:,:), TFORCE(:,:,:), QSOURCE(:,:,:)
  real, allocatable, save, public :: FLOWX(:,:,:), FLOWT(:,:,:), FLOWR(:,:,:), ROAVG_CELL(:,:,:)
  ! the Spalart-Allmaras variable: zero, or read from the restart file
  real, allocatable, save, public :: TRANS_DYN_VIS(:,:,:)

  ! ---- multigrid block time steps ------------------------------------------
  real, allocatable, save, public :: STEP1(:,:,:), STEP2(:,:,:), STEPSBK(:)
  ! wall distances and mixing lengths (SET_XLENGTH)
  real, allocatable, save, public :: DWALLSQ(:,:,:), XLENGTH(:,:,:), FST_RAT(:,:,:), XLIMIT(:)

contains

  ! INTP, mf:7801-7836: four-point Lagrange interpolation in a table.
  subroutine intp(n, xn, yn, xa, y)
    integer, intent(in) :: n
    real, intent(in) :: xn(n), yn(n), xa
    real, intent(out) :: y
    real :: span, co, temp
    integer :: l, nm, l1, l2
    span = xn(n) - xn(1)
    y = 0.
    l = 1
    nm = n
    if (n >= 4) then
      nm = 4
      do
        if (span > 0.0 .and. xa < xn(l)) exit
        if (span < 0.0 .and. xa > xn(l)) exit
        if (l == n) exit
        l = l + 1
      end do
      if (l == n .and. .not. ((span > 0.0 .and. xa < xn(l)) .or. (span < 0.0 .and. xa > xn(l)))) then
        l = n - 3
      else if (l <= 2) then
        l = 1
      else if (l == n) then
        l = n - 3
      else
        l = l - 2
      end if
    end if
    do l1 = 1, nm
      co = 1
      do l2 = 1, nm
        if (l1 == l2) then
          temp = 1
        else
          temp = (xa - xn(l + l2 - 1)) / (xn(l + l1 - 1) - xn(l + l2 - 1))
        end if
        co = co * temp
      end do
      y = y + co * yn(l + l1 - 1)
    end do
  end subroutine intp

  ! ---------------------------------------------------------------------------
  ! list-directed value lines with the reference's defaults on error
  subroutine skip(u)
    integer, intent(in) :: u
    character(len=256) :: line
    read (u, '(A)') line
  end subroutine skip

  logical function fail(msg)
    character(len=*), intent(in) :: msg
    mh_host_err = msg
    fail = .true.
  end function fail

  ! ---------------------------------------------------------------------------
  ! NEW_READIN, mf:148-2158 (the options listed in the module header)
  logical function read_deck(u) result(bad)
    integer, intent(in) :: u
    integer, parameter :: JCAP = 20000, KCAP = 400
    integer :: ios, nr, k, j, i, n, j1, j2, jmrow, jlerow, jterow, nblades_in_row, new_grid, jrow
    integer :: jroths, jrothe, jrotts, jrotte, nsecs_row, insurf, if_design, if_stagger, if_lean, jmidrw
    integer :: kin_mid, iout(20), n_angles, j_mid
    real :: fracn_span(KCAP), angl_up(KCAP), angl_dwn1(KCAP), angl_dwn2(KCAP), qspan(KCAP), xdif, rdif
    integer, allocatable :: kout(:)
    real :: rpmrow, rpmhub, puprow, plerow, pterow, pdnrow, fac1, xshift, fac2, tshift, fac3, fac4, rshift
    real :: xd, rd, srange, slim, xspan, rspan, f_span, xarg, rarg, frrat, frmax, frev
    real :: sdist(JCAP), xint1(KCAP), xint2(KCAP), xint3(KCAP), xint4(KCAP)
    real :: xnewhub(JCAP), rnewhub(JCAP), xnewtip(JCAP), rnewtip(JCAP)
    character(len=72) :: title, rowtyp
    integer, allocatable :: t_ind(:), t_indle(:), t_indte(:), t_indmix(:), t_indmid(:), t_nblade(:), t_nrow(:)
    real, allocatable :: t_wrad(:), t_whub(:), t_wtip(:), t_pguess(:)

    bad = .false.
    read (u, '(A)') title
    call skip(u)
    read (u, *, iostat=ios) CP, GA
    if (CP < 0.0) then
      bad = fail('real-gas properties (CP<0) are not supported'); return
    end if
    IFGAS = 0
    ITIMST = 3
    call skip(u)
    read (u, *, iostat=ios) ITIMST
    if (ITIMST == 3 .or. ITIMST == 5 .or. ITIMST == 6) then
      F1 = 2.0000; F2 = -1.000; F3 = 0.00; F2EFF = -1.0; NRSMTH = 0; RSMTH = 0.40
    end if
    if (ITIMST == -3 .or. ITIMST == -5 .or. ITIMST == -6) then
      F1 = 2.0000; F2 = -1.65; F3 = -0.65; F2EFF = -1.0; NRSMTH = 1; RSMTH = 0.40
      ITIMST = abs(ITIMST)
    end if
    if (ITIMST == 4 .or. ITIMST == -4) then
      read (u, *) F1, F2EFF, F3, RSMTH, NRSMTH
      if (F2EFF > 0.0) F2EFF = -F2EFF
      if (F3 > 0.0) F3 = -F3
      F2 = F2EFF * (1.0 - F3)
      ITIMST = 3
    end if
    if (ITIMST >= 5) then
      bad = fail('artificial compressibility (ITIMST 5/6) is not supported'); return
    end if
    CFL = 0.4; DAMPIN = 10.0; MACHLIM = 2.0; F_PDOWN = 0.0
    call skip(u)
    read (u, *, iostat=ios) CFL, DAMPIN, MACHLIM, F_PDOWN
    IF_RESTART = 0
    call skip(u)
    read (u, *, iostat=ios) IF_RESTART
    NSTEPS_MAX = 5000; CONLIM = 0.005
    call skip(u)
    read (u, *, iostat=ios) NSTEPS_MAX, CONLIM
    SFXIN = 0.005; SFTIN = 0.005; FAC_4TH = 0.8; NCHANGE = NSTEPS_MAX/4
    call skip(u)
    read (u, *, iostat=ios) SFXIN, SFTIN, FAC_4TH, NCHANGE
    if (IF_RESTART /= 0) NCHANGE = 100   ! mf:382
    call skip(u)
    read (u, *) NROWS
    call skip(u)
    read (u, *) IM, KM
    if (IM < 4 .or. KM < 4) then
      bad = fail('throughflow (IM=2) and Q3D (KM=2) are not supported'); return
    end if
    IMM1 = IM - 1; IMM2 = IM - 2; KMM1 = KM - 1; KMM2 = KM - 2
    allocate (FP(IM), FR(KM), FU(IM), FD(IM), FKU(KM), FKD(KM), FSPAN(KM))
    FP = 0.; FR = 0.
    call skip(u)
    read (u, *) (FP(i), i = 1, IMM1)
    call skip(u)
    read (u, *) (FR(k), k = 1, KMM1)
    IR = 3; JR = 3; KR = 3; IRBB = 9; JRBB = 9; KRBB = 9
    call skip(u)
    read (u, *, iostat=ios) IR, JR, KR, IRBB, JRBB, KRBB
    FBLK1 = 0.4; FBLK2 = 0.2; FBLK3 = 0.1
    call skip(u)
    read (u, *, iostat=ios) FBLK1, FBLK2, FBLK3
    IFMIX = 1
    call skip(u)
    read (u, *, iostat=ios) IFMIX
    if (IFMIX /= 0) then
      RFMIX = 0.025; FSMTHB = 1.0; FEXTRAP = 0.80; FANGLE = 0.80
      call skip(u)
      read (u, *, iostat=ios) RFMIX, FEXTRAP, FSMTHB, FANGLE
    end if
    IFCOOL = 0; IFBLEED = 0; IF_ROUGH = 0
    call skip(u)
    read (u, *, iostat=ios) IFCOOL, IFBLEED, IF_ROUGH
    if (IFCOOL /= 0 .or. IFBLEED /= 0) then
      bad = fail('cooling and bleed flows are not supported'); return
    end if
    call skip(u)
    read (u, *) NSECS_IN
    IN_PRESS = 0; IN_VTAN = 0; IN_VR = 1; IN_FLOW = 0; IF_REPEAT = 0; RFIN = 0.1
    call skip(u)
    read (u, *, iostat=ios) IN_PRESS, IN_VTAN, IN_VR, IN_FLOW, IF_REPEAT, RFIN
    IPOUT = 1; SFEXIT = 0.0; NSFEXIT = 0
    call skip(u)
    read (u, *, iostat=ios) IPOUT, SFEXIT, NSFEXIT
    PLATE_LOSS = 0.0; THROTTLE_EXIT = 0.0
    call skip(u)
    read (u, *, iostat=ios) PLATE_LOSS, THROTTLE_EXIT
    if (THROTTLE_EXIT > 0.001 .or. PLATE_LOSS > 0.001) then
      bad = fail('exit throttle and plate loss are not supported'); return
    end if
    ! mass-flow forcing target and relaxation factor, mf:567-572
    FLOWIN = 0.0; RFLOW = 0.0
    if (IN_FLOW /= 0) then
      call skip(u)
      read (u, *) FLOWIN, RFLOW
    end if
    ! the repeating-stage factors, mf:575-585
    NINMOD = 0; RFINBC = 0.0
    if (IF_REPEAT /= 0) then
      NINMOD = 10; RFINBC = 0.025
      call skip(u)
      read (u, *, iostat=ios) NINMOD, RFINBC
    end if
    ILOS = 100; NLOS = 5; IBOUND = 0
    call skip(u)
    read (u, *, iostat=ios) ILOS, NLOS, IBOUND
    REYNO = 500000.0; RF_VIS = 0.5; FTRANS = 0.0001; PRANDTL = 1.0; YPLUSWALL = 0.0; TURBVIS_LIM = 3000.0
    if (ILOS /= 0) then
      call skip(u)
      read (u, *, iostat=ios) REYNO, RF_VIS, FTRANS, TURBVIS_LIM, PRANDTL, YPLUSWALL
    end if
    if (ILOS >= 200) then
      FAC_STMIX = 0.0; FAC_ST0 = 1.0; FAC_ST1 = 1.0; FAC_ST2 = 1.0; FAC_ST3 = 1.0; FAC_SFVIS = 2.0
      FAC_VORT = 0.0; FAC_PGRAD = 0.0
      call skip(u)
      read (u, *, iostat=ios) FAC_STMIX, FAC_ST0, FAC_ST1, FAC_ST2, FAC_ST3, FAC_SFVIS, FAC_VORT, FAC_PGRAD
    end if
    if (ILOS /= 0) then
      YPLAM = 5.0; YPTURB = 25.0
      call skip(u)
      read (u, *, iostat=ios) YPLAM, YPTURB
    end if
    ISHIFT = 2; NEXTRAP_LE = 10; NEXTRAP_TE = 10
    call skip(u)
    read (u, *, iostat=ios) ISHIFT, NEXTRAP_LE, NEXTRAP_TE
    if (ISHIFT < 0 .or. ISHIFT > 4) then
      bad = fail('ISHIFT must be 0 to 4'); return
    end if
    allocate (NSTAGE(NROWS + 1))
    do n = 1, NROWS
      NSTAGE(n) = 1 + (n - 1)/2
    end do
    call skip(u)
    read (u, *, iostat=ios) (NSTAGE(n), n = 1, NROWS)
    NOUT = NSTEPS_MAX + 10
    call skip(u)
    read (u, *, iostat=ios) (NOUT(n), n = 1, 5)
    call skip(u)                     ! output markers IOUT, KOUT (used only for printing)
    read (u, *, iostat=ios) (iout(i), i = 1, 13)
    call skip(u)
    allocate (kout(KM))
    read (u, *, iostat=ios) (kout(k), k = 1, KM)

    allocate (JSTART(NROWS + 1), JMIX(NROWS + 1), JLE(NROWS + 1), JTE(NROWS + 1), JLED(NROWS + 1))
    allocate (KTIPS(NROWS + 1), KTIPE(NROWS + 1), KS1(NROWS + 1), KS2(NROWS + 1), IF_CUSP(NROWS))
    allocate (IF_ANGLES(NROWS), FRACTIP1(NROWS), FRACTIP2(NROWS), FTHICK(NROWS, KM), CHORD(NROWS))
    allocate (JTRAN_I1(NROWS), JTRAN_IM(NROWS), JTRAN_K1(NROWS), JTRAN_KM(NROWS))
    allocate (ICUSP(NROWS), LCUSP(NROWS), LCUSPUP(NROWS))
    allocate (NUP_I1(NROWS), NUP_IM(NROWS), N_WAKE(NROWS), SEP_THIK(NROWS), SEP_DRAG(NROWS))
    NUP_I1 = 0; NUP_IM = 0; N_WAKE = 0; SEP_THIK = 0.0; SEP_DRAG = 0.0
    allocate (BETAUP(NROWS, max(NSECS_IN, 2)), BETADWN1(NROWS, max(NSECS_IN, 2)), BETADWN2(NROWS, max(NSECS_IN, 2)))
    ICUSP = 0; LCUSP = 0; LCUSPUP = 0; BETAUP = 0.0; BETADWN1 = 0.0; BETADWN2 = 0.0
    allocate (t_ind(JCAP), t_indle(JCAP), t_indte(JCAP), t_indmix(JCAP), t_indmid(JCAP), t_nblade(JCAP))
    allocate (t_nrow(JCAP), t_wrad(JCAP), t_whub(JCAP), t_wtip(JCAP), t_pguess(JCAP))
    allocate (XSURF(JCAP, max(NSECS_IN, 2)), RSURF(JCAP, max(NSECS_IN, 2)))
    allocate (RT_UPP(JCAP, max(NSECS_IN, KM)), RT_THICK(JCAP, max(NSECS_IN, KM)))
    JSTART = 0; JMIX = 0; JLE = 0; JTE = 0; JLED = 0; KTIPS = 0; KTIPE = 0; KS1 = 0; KS2 = 0
    t_indle = 0
    j1 = 1
    do nr = 1, NROWS
      call skip(u)
      call skip(u)
      read (u, '(A)') rowtyp
      call skip(u)
      read (u, *) nblades_in_row
      call skip(u)
      read (u, *) jmrow, jlerow, jterow
      if (j1 + jmrow > JCAP) then
        bad = fail('too many J points'); return
      end if
      call skip(u)
      read (u, *) KTIPS(nr), KTIPE(nr)
      if (KTIPS(nr) <= 0) KTIPE(nr) = 0
      if (KTIPS(nr) < 0) then
        bad = fail('shroud leakage (KTIPS<0) is not supported'); return
      end if
      if (KTIPS(nr) > 0) then
        call skip(u)
        read (u, *) FRACTIP1(nr), FRACTIP2(nr)
      end if
      FTHICK(nr, :) = 1.0
      if (KTIPS(nr) > 0 .and. KTIPE(nr) > 0) then
        call skip(u)
        read (u, *) (FTHICK(nr, k), k = 1, KM)
      end if
      call skip(u)
      read (u, *) JTRAN_I1(nr), JTRAN_IM(nr), JTRAN_K1(nr), JTRAN_KM(nr)
      call skip(u)
      read (u, *) new_grid
      if (new_grid /= 0) then
        bad = fail('NEWGRID is not supported'); return
      end if
      call skip(u)
      read (u, *) rpmrow, rpmhub
      call skip(u)
      read (u, *) jroths, jrothe, jrotts, jrotte
      if (jroths == 0 .or. jroths > jmrow) jroths = jmrow
      if (jrothe == 0 .or. jrothe > jmrow) jrothe = jmrow
      if (jrotts == 0 .or. jrotts > jmrow) jrotts = jmrow
      if (jrotte == 0 .or. jrotte > jmrow) jrotte = jmrow
      call skip(u)
      read (u, *) puprow, plerow, pterow, pdnrow
      call skip(u)
      nsecs_row = NSECS_IN
      insurf = 0
      read (u, *, iostat=ios) nsecs_row, insurf
      if (NSECS_IN == 1) nsecs_row = 1
      if (NSECS_IN == 1) NSECS_IN = 2
      if (nsecs_row == 1) nsecs_row = 2
      if (insurf /= 0 .or. nsecs_row /= NSECS_IN) then
        bad = fail('INSURF>0 or NSECS_ROW/=NSECS_IN is not supported'); return
      end if
      IF_CUSP(nr) = 0; IF_ANGLES(nr) = 0
      call skip(u)
      read (u, *, iostat=ios) IF_CUSP(nr), IF_ANGLES(nr)
      if (IF_CUSP(nr) < 0 .or. IF_CUSP(nr) > 2 .or. IF_ANGLES(nr) < 0) then
        bad = fail('IF_CUSP must be 0, 1 or 2'); return
      end if
      ! cusp shape, mf:980-984
      if (IF_CUSP(nr) == 1) read (u, *) ICUSP(nr), LCUSP(nr), LCUSPUP(nr)
      ! the trailing-edge separation force, mf:993-1004
      if (IF_CUSP(nr) == 2) then
        if (ISHIFT >= 2) then
          ! the re-gridded gap would form a new cusp of zero length there
          bad = fail('IF_CUSP = 2 with ISHIFT >= 2 (gap re-gridding) is not supported'); return
        end if
        read (u, *) NUP_I1(nr), NUP_IM(nr), N_WAKE(nr), SEP_THIK(nr), SEP_DRAG(nr)
      end if
      j2 = j1 + jmrow - 1
      do k = 1, nsecs_row
        if_design = 0; if_stagger = 0; if_lean = 0
        call skip(u)
        call skip(u)
        read (u, *, iostat=ios) if_design, if_stagger, if_lean
        if (if_design /= 0 .or. if_stagger /= 0 .or. if_lean /= 0) then
          bad = fail('section redesign, restagger and lean are not supported'); return
        end if
        read (u, *) fac1, xshift
        read (u, *) (XSURF(j, k), j = j1, j2)
        read (u, *) fac2, tshift
        read (u, *) (RT_UPP(j, k), j = j1, j2)
        read (u, *) fac3
        read (u, *) (RT_THICK(j, k), j = j1, j2)
        read (u, *) fac4, rshift
        read (u, *) (RSURF(j, k), j = j1, j2)
        ! mf:1446-1470: move points closer than 1e-4 of the row length
        sdist(j1) = 0.0
        do j = j1 + 1, j2
          xd = XSURF(j, k) - XSURF(j - 1, k)
          rd = RSURF(j, k) - RSURF(j - 1, k)
          sdist(j) = sdist(j - 1) + sqrt(xd*xd + rd*rd)
        end do
        srange = sdist(j2) - sdist(j1)
        slim = 0.0001*srange
        do j = j1 + 2, j2
          if ((sdist(j) - sdist(j - 1)) < slim) then
            xd = slim*(XSURF(j - 2, k) - XSURF(j - 1, k))/(sdist(j - 2) - sdist(j - 1))
            rd = slim*(RSURF(j - 2, k) - RSURF(j - 1, k))/(sdist(j - 2) - sdist(j - 1))
            XSURF(j, k) = XSURF(j - 1, k) + xd
            RSURF(j, k) = RSURF(j - 1, k) + rd
          end if
        end do
        do j = j1, j2
          XSURF(j, k) = fac1*(xshift + XSURF(j, k))
          RT_THICK(j, k) = fac3*RT_THICK(j, k)
          RSURF(j, k) = fac4*(rshift + RSURF(j, k))
          RT_UPP(j, k) = fac2*(tshift + RT_UPP(j, k))
        end do
      end do
      ! hub and casing lines from the first and last sections, mf:1650-1700
      do j = j1, j2
        xnewhub(j) = XSURF(j, 1)
        rnewhub(j) = RSURF(j, 1)
        xnewtip(j) = XSURF(j, nsecs_row)
        rnewtip(j) = RSURF(j, nsecs_row)
      end do
      ! NSECS_IN sections equally spaced between hub and casing, mf:1700-1740
      do j = j1, j2
        xspan = xnewtip(j) - xnewhub(j)
        rspan = rnewtip(j) - rnewhub(j)
        do k = 1, nsecs_row
          xint1(k) = XSURF(j, k)
          xint2(k) = RSURF(j, k)
          xint3(k) = RT_UPP(j, k)
          xint4(k) = RT_THICK(j, k)
        end do
        if (abs(xspan) > abs(rspan)) then
          do k = 1, NSECS_IN
            f_span = float(k - 1)/float(NSECS_IN - 1)
            xarg = xnewhub(j) + f_span*(xnewtip(j) - xnewhub(j))
            XSURF(j, k) = xarg
            call intp(nsecs_row, xint1, xint2, xarg, RSURF(j, k))
            call intp(nsecs_row, xint1, xint3, xarg, RT_UPP(j, k))
            call intp(nsecs_row, xint1, xint4, xarg, RT_THICK(j, k))
          end do
        else
          do k = 1, NSECS_IN
            f_span = float(k - 1)/float(NSECS_IN - 1)
            rarg = rnewhub(j) + f_span*(rnewtip(j) - rnewhub(j))
            RSURF(j, k) = rarg
            call intp(nsecs_row, xint2, xint1, rarg, XSURF(j, k))
            call intp(nsecs_row, xint2, xint3, rarg, RT_UPP(j, k))
            call intp(nsecs_row, xint2, xint4, rarg, RT_THICK(j, k))
          end do
        end if
      end do
      ! upstream and downstream grid angles on the sections, mf:1342-1374
      if (IF_ANGLES(nr) > 0) then
        call skip(u)
        read (u, *) n_angles
        read (u, *) (fracn_span(n), n = 1, n_angles)
        read (u, *) (angl_up(n), n = 1, n_angles)
        read (u, *) (angl_dwn1(n), n = 1, n_angles)
        read (u, *) (angl_dwn2(n), n = 1, n_angles)
        qspan(1) = 0.0
        j_mid = j1 + (jlerow + jterow)/2
        do k = 2, NSECS_IN
          xdif = XSURF(j_mid, k) - XSURF(j_mid, k - 1)
          rdif = RSURF(j_mid, k) - RSURF(j_mid, k - 1)
          qspan(k) = qspan(k - 1) + sqrt(xdif*xdif + rdif*rdif)
        end do
        do k = 1, NSECS_IN
          qspan(k) = qspan(k)/qspan(NSECS_IN)
        end do
        do k = 1, NSECS_IN
          call intp(n_angles, fracn_span, angl_up, qspan(k), BETAUP(nr, k))
          call intp(n_angles, fracn_span, angl_dwn1, qspan(k), BETADWN1(nr, k))
          call intp(n_angles, fracn_span, angl_dwn2, qspan(k), BETADWN2(nr, k))
        end do
      end if
      ! markers, rotation and pressure guess for the row, mf:1765-1840
      JLE(nr) = j1 + jlerow - 1
      JTE(nr) = j1 + jterow - 1
      j2 = j1 - 1 + jmrow
      jrow = 0
      do j = j1, j2
        jrow = jrow + 1
        t_wrad(j) = rpmrow*PI/30.
        t_whub(j) = 0.0
        if (jrow > jroths .and. jrow < jrothe) t_whub(j) = rpmhub*PI/30.
        t_wtip(j) = 0.0
        if (jrow > jrotts .and. jrow < jrotte) t_wtip(j) = t_wrad(j)
        if (KTIPS(nr) <= 0) then
          KS1(nr) = 1
          KS2(nr) = KMM1
        end if
        if (KTIPS(nr) == 1) then
          KS1(nr) = KTIPE(nr)
          KS2(nr) = KMM1
        end if
        if (KTIPS(nr) > 1) then
          KS1(nr) = 1
          KS2(nr) = KTIPS(nr) - 1
          KTIPE(nr) = KM
        end if
        t_ind(j) = 0
        t_indle(j) = 0
        t_indte(j) = 0
        t_indmix(j) = 0
        t_indmid(j) = 0
        t_nblade(j) = nblades_in_row
        t_nrow(j) = nr
        jmidrw = int(0.5*(jlerow + jterow))
        if (jrow > jlerow .and. jrow <= jterow) t_ind(j) = 1
        if (jrow == jlerow) t_indle(j) = 1
        if (jrow == jlerow) JLED(nr) = j
        if (jrow == jterow) t_indte(j) = 1
        if (jrow == jlerow .and. j >= 2) t_indle(j - 1) = -1
        if (jrow == jlerow .and. j >= 3) t_indle(j - 2) = -2
        if (jrow == jmidrw) t_indmid(j) = 1
        if (jrow <= jlerow) t_pguess(j) = puprow + (plerow - puprow)*(jrow - 1)/(jlerow - 1)
        if (jrow > jlerow .and. jrow <= jterow) &
          t_pguess(j) = plerow + (pterow - plerow)*(jrow - jlerow)/(jterow - jlerow)
        if (jrow > jterow) t_pguess(j) = pterow + (pdnrow - pterow)*(jrow - jterow)/(jmrow - jterow)
      end do
      t_indmix(j2) = 1
      JMIX(nr) = j2
      JSTART(nr) = j1
      j1 = j2 + 1
    end do
    JM = j2
    JMM1 = JM - 1
    JMM2 = JM - 2
    t_indmix(JM) = 0
    JSTART(NROWS + 1) = JM + 1
    allocate (IND(JM), INDLE(JM), INDTE(JM), INDMIX(JM), INDMID(JM), NBLADE(JM), NROW(JM))
    allocate (WRAD(JM), WHUB(JM), WTIP(JM), PGUESS(JM))
    IND = t_ind(1:JM); INDLE = t_indle(1:JM); INDTE = t_indte(1:JM); INDMIX = t_indmix(1:JM)
    INDMID = t_indmid(1:JM); NBLADE = t_nblade(1:JM); NROW = t_nrow(1:JM); WRAD = t_wrad(1:JM)
    WHUB = t_whub(1:JM); WTIP = t_wtip(1:JM); PGUESS = t_pguess(1:JM)
    if (ISHIFT /= 0) then
      bad = grid_gaps()
      if (bad) return
    end if

    ! inlet and exit boundary data, mf:1838-1940
    call skip(u)
    call skip(u)
    read (u, *) KIN
    KIN_MID = int(0.5*KIN)
    allocate (FR_IN(max(KIN, KM)), PO1(max(KIN, KM)), TO1(max(KIN, KM)), VTIN(max(KIN, KM)))
    allocate (VM1(max(KIN, KM)), BS(max(KIN, KM)), BR(max(KIN, KM)), PD(max(KIN, KM)))
    FR_IN = 0.; PD = 0.
    call skip(u)
    read (u, *) (FR_IN(k), k = 1, KIN - 1)
    call skip(u)
    read (u, *) (PO1(k), k = 1, KIN)
    PO_IN_MID = PO1(KIN_MID)
    if (IPOUT == 3) then
      call skip(u)
      read (u, *) (PD(k), k = 1, KIN)
    end if
    call skip(u)
    read (u, *) (TO1(k), k = 1, KIN)
    call skip(u)
    read (u, *) (VTIN(k), k = 1, KIN)
    VT_IN_MID = VTIN(KIN_MID)
    call skip(u)
    read (u, *) (VM1(k), k = 1, KIN)
    call skip(u)
    read (u, *) (BS(k), k = 1, KIN)
    call skip(u)
    read (u, *) (BR(k), k = 1, KIN)
    call skip(u)
    read (u, *) PDOWN_HUB, PDOWN_TIP
    ! geometric series when the third spacing is zero, mf:1955-2030
    if (.not. (FR(3) > 0.00001)) then
      frrat = FR(1); frmax = FR(2); FR(1) = 1.0; FR(KM) = 0.0
      do k = 2, KMM1
        FR(k) = FR(k - 1)*frrat
      end do
      do k = 1, KMM1
        frev = FR(KM - k)
        if (frev < FR(k)) FR(k) = frev
        if (FR(k) > frmax) FR(k) = frmax
      end do
    end if
    if (KIN == 2) then
      FR_IN(1) = 1.0; FR_IN(2) = 0.0
    else if (.not. (FR_IN(3) > 0.00001)) then
      frrat = FR_IN(1); frmax = FR_IN(2); FR_IN(1) = 1.0; FR_IN(KIN) = 0.0
      do k = 2, KIN - 1
        FR_IN(k) = FR_IN(k - 1)*frrat
      end do
      do k = 1, KIN - 1
        frev = FR_IN(KIN - k)
        if (frev < FR_IN(k)) FR_IN(k) = frev
        if (FR_IN(k) > frmax) FR_IN(k) = frmax
      end do
    end if
    if (.not. (FP(3) > 0.00001)) then
      frrat = FP(1); frmax = FP(2); FP(1) = 1.0; FP(IM) = 0.0
      do i = 2, IMM1
        FP(i) = FP(i - 1)*frrat
      end do
      do i = 1, IMM1
        frev = FP(IM - i)
        if (frev < FP(i)) FP(i) = frev
        if (FP(i) > frmax) FP(i) = frmax
      end do
    end if
    call inpint()
    ! viscous-model limits, mf:2090-2125
    allocate (XLLIM_I1(NROWS), XLLIM_IM(NROWS), XLLIM_K1(NROWS), XLLIM_KM(NROWS), XLLIM_DWN(NROWS))
    allocate (XLLIM_UP(NROWS), XLLIM_IN(NROWS), XLLIM_LE(NROWS), XLLIM_TE(NROWS), XLLIM_DN(NROWS))
    allocate (FSTURB(NROWS), TURBVIS_DAMP(NROWS), ROUGH_H(NROWS), ROUGH_T(NROWS), ROUGH_L(NROWS), ROUGH_U(NROWS))
    XLLIM_I1 = 0.03; XLLIM_IM = 0.03; XLLIM_K1 = 0.03; XLLIM_KM = 0.03; XLLIM_DWN = 0.03; XLLIM_UP = 0.02
    XLLIM_IN = 0.02; XLLIM_LE = 0.03; XLLIM_TE = 0.04; XLLIM_DN = 0.05; FSTURB = 1.0; TURBVIS_DAMP = 0.5
    FACMIXUP = 0.0
    NMIXUP = 0
    if (ILOS /= 0) then
      FACMIXUP = 2.0
      NMIXUP = 1000
      call skip(u)
      do n = 1, NROWS
        if (ILOS == 10) read (u, *, iostat=ios) XLLIM_I1(n), XLLIM_IM(n), XLLIM_K1(n), XLLIM_KM(n), &
          XLLIM_DWN(n), XLLIM_UP(n)
        if (ILOS >= 100) read (u, *, iostat=ios) XLLIM_IN(n), XLLIM_LE(n), XLLIM_TE(n), XLLIM_DN(n), &
          FSTURB(n), TURBVIS_DAMP(n)
      end do
      call skip(u)
      read (u, *, iostat=ios) FACMIXUP, NMIXUP
      if (ios == 0) then    ! mf:3490-3492, skipped by a read error
        if (FACMIXUP < 1.0) FACMIXUP = 1.0
        if (IF_RESTART /= 0) FACMIXUP = 1.0
      end if
    end if
    CFWALL = 0.0
    if (YPLUSWALL > 5.0) CFWALL = 1/(YPLUSWALL*YPLUSWALL)
    ROUGH_H = 0.0; ROUGH_T = 0.0; ROUGH_L = 0.0; ROUGH_U = 0.0
    if (IF_ROUGH > 0) then
      call skip(u)
      do n = 1, NROWS
        read (u, *) ROUGH_H(n), ROUGH_T(n), ROUGH_L(n), ROUGH_U(n)
        ROUGH_H(n) = ROUGH_H(n)*1.0E-06
        ROUGH_T(n) = ROUGH_T(n)*1.0E-06
        ROUGH_L(n) = ROUGH_L(n)*1.0E-06
        ROUGH_U(n) = ROUGH_U(n)*1.0E-06
      end do
    end if
  end function read_deck

  ! ---------------------------------------------------------------------------
  ! The grid between blade rows, NEW_READIN mf:1546-1837.  ISHIFT = 1 moves
  ! each row axially so the hub X spacing runs on across the mixing plane;
  ! ISHIFT = 2 to 4 re-grid every gap (upstream of the first row, between
  ! rows, downstream of the last) on each section: the gap's meridional
  ! curve (kept, straightened, or straightened away from hub and casing),
  ! smoothed, then laid out by GRID_UP / GRID_DOWN.
  logical function grid_gaps() result(bad)
    real :: sm(JM), xint(JM), rint(JM), sdist(JM)
    real :: xshift, diff1, diff2, dlast, xgap, rgap, gap, xd, rd, proj, dist, sfgap, smid, ste, sle
    real :: anglup, angldwn1, angldwn2
    integer :: j, k, ngap, jst, jend, jmid, nrw, nrwm1, npoints, ngaps, ns, nn
    bad = .false.
    if (ISHIFT == 1) then
      xshift = 0.0
      do j = 2, JM
        if (INDMIX(j - 1) == 1) then
          diff1 = XSURF(j - 1, 1) - XSURF(j - 2, 1)
          diff2 = XSURF(j + 1, 1) - XSURF(j, 1)
          xshift = XSURF(j - 1, 1) + 0.01*(diff1 + diff2) - XSURF(j, 1)
        end if
        do k = 1, NSECS_IN
          XSURF(j, k) = XSURF(j, k) + xshift
        end do
      end do
      return
    end if
    jmid = 0
    do ngap = 1, NROWS + 1
      if (ngap == 1) then
        nrw = 1; jst = 1; jend = JLE(1)
      end if
      if (ngap > 1 .and. ngap /= NROWS + 1) then
        nrw = ngap; jst = JTE(nrw - 1); jend = JLE(nrw); jmid = JMIX(nrw - 1)
      end if
      if (ngap == NROWS + 1) then
        nrw = NROWS + 1; jst = JTE(NROWS); jend = JM
      end if
      nrwm1 = nrw - 1
      do k = 1, NSECS_IN
        sm(1) = 0.0
        do j = 2, JM
          xd = XSURF(j, k) - XSURF(j - 1, k)
          rd = RSURF(j, k) - RSURF(j - 1, k)
          sm(j) = sm(j - 1) + sqrt(xd*xd + rd*rd)
        end do
        ! the points of the gap's meridional curve
        dlast = -1.0
        npoints = 1
        xgap = XSURF(jend, k) - XSURF(jst, k)
        rgap = RSURF(jend, k) - RSURF(jst, k)
        gap = sqrt(xgap*xgap + rgap*rgap)
        do j = jst, jend
          xd = XSURF(j, k) - XSURF(jst, k)
          rd = RSURF(j, k) - RSURF(jst, k)
          proj = (xd*xgap + rd*rgap)
          if (ISHIFT >= 3) proj = proj/gap
          dist = sqrt(xd*xd + rd*rd)
          if (proj < 0.0) cycle
          if (dist > (1.00001*gap)) then
            xint(npoints) = XSURF(jend, k)
            rint(npoints) = RSURF(jend, k)
            npoints = npoints + 1
            exit
          end if
          if (dist < dlast) cycle
          dlast = dist
          if (ISHIFT == 3 .or. (ISHIFT == 4 .and. k /= 1 .and. k /= NSECS_IN)) then
            xint(npoints) = XSURF(jst, k) + proj*xgap/gap
            rint(npoints) = RSURF(jst, k) + proj*rgap/gap
          else
            xint(npoints) = XSURF(j, k)
            rint(npoints) = RSURF(j, k)
          end if
          npoints = npoints + 1
        end do
        ! ten smoothing passes along the curve, then its arc length
        ngaps = npoints - 1
        sfgap = 0.2
        do ns = 1, 10
          do nn = 2, ngaps - 1
            xint(nn) = (1. - sfgap)*xint(nn) + sfgap*0.5*(xint(nn - 1) + xint(nn + 1))
            rint(nn) = (1. - sfgap)*rint(nn) + sfgap*0.5*(rint(nn - 1) + rint(nn + 1))
          end do
        end do
        sdist(1) = sm(jst)
        do nn = 2, ngaps
          xd = xint(nn) - xint(nn - 1)
          rd = rint(nn) - rint(nn - 1)
          sdist(nn) = sdist(nn - 1) + sqrt(xd*xd + rd*rd)
        end do
        smid = 0.5*(sdist(1) + sdist(ngaps))
        ste = sdist(1)
        sle = sdist(ngaps)
        anglup = 0.0
        if (nrw <= NROWS) anglup = BETAUP(nrw, k)
        if (nrw == 1) then
          angldwn1 = BETADWN1(1, k); angldwn2 = BETADWN2(1, k)
        else
          angldwn1 = BETADWN1(nrwm1, k); angldwn2 = BETADWN2(nrwm1, k)
        end if
        if (nrw == NROWS + 1) then
          angldwn1 = BETADWN1(NROWS, k); angldwn2 = BETADWN2(NROWS, k)
        end if
        if (nrw == 1) then
          bad = grid_up(k, jst, jend, ste, sle, ngaps, sdist, xint, rint, sm, NEXTRAP_LE, anglup, IF_ANGLES(1))
        else if (nrw /= NROWS + 1) then
          bad = grid_down(k, jst, jmid, ste, smid, ngaps, sdist, xint, rint, sm, NEXTRAP_TE, angldwn1, angldwn2, &
                          IF_CUSP(nrwm1), ICUSP(nrwm1), LCUSP(nrwm1), LCUSPUP(nrwm1), IF_ANGLES(nrwm1))
          if (bad) return
          bad = grid_up(k, jmid + 1, jend, smid, sle, ngaps, sdist, xint, rint, sm, NEXTRAP_LE, anglup, &
                        IF_ANGLES(nrw))
        else
          bad = grid_down(k, jst, jend, ste, sle, ngaps, sdist, xint, rint, sm, NEXTRAP_TE, angldwn1, angldwn2, &
                          IF_CUSP(NROWS), ICUSP(NROWS), LCUSP(NROWS), LCUSPUP(NROWS), IF_ANGLES(NROWS))
        end if
        if (bad) return
      end do
    end do
  end function grid_gaps

  ! SOLVE, mf:9513-9543: the ratio RAT of a geometric progression of NPUP
  ! spacings, the first DX1, that adds up to XUP.
  subroutine solve(npup, xup, dx1, rat)
    integer, intent(in) :: npup
    real, intent(in) :: xup, dx1
    real, intent(out) :: rat
    real :: rup, rhs, rnew, diff
    integer :: m
    rup = xup/dx1
    rat = 1.2
    if (rup < npup) rat = 0.9
    m = 1
    do
      rhs = (rat - 1.)*rup
      if (rup >= npup) rnew = (rhs + 1.)**(1./npup)
      if (rup < npup) rnew = (rat**npup - 1.)/rup + 1.
      diff = abs(rnew/rat - 1.0)
      rat = rnew
      m = m + 1
      if (diff < 0.00001) exit
      if (m > 250) exit
    end do
  end subroutine solve

  ! GRID_UP, mf:9433-9508: J1..J2 upstream of a leading edge at J2, spaced
  ! in a geometric progression from the first blade spacing, placed on the
  ! smoothed gap curve; the blade's camber line is run on upstream at the
  ! slope of its first NEXTRAP points (or the input angle).
  logical function grid_up(k, j1, j2, s1, s2, ngap, sdist, xint, rint, sm, nextrap, beta_up, ifangles) result(bad)
    integer, intent(in) :: k, j1, j2, ngap, nextrap, ifangles
    real, intent(in) :: s1, s2, sdist(:), xint(:), rint(:), sm(:), beta_up
    real :: snew(JM), ansx(JM), ansr(JM), th_upp(JM), th_thick(JM)
    real :: sdiff, r1, rat, xd, slopet, rtheta_le
    integer :: j, jsub, nint
    bad = .false.
    if (j2 + nextrap > JM .or. j2 - j1 < 1) then
      bad = fail('GRID_UP: the row has fewer than NEXTRAP_LE points after its leading edge'); return
    end if
    snew(j1) = s1
    snew(j2) = s2
    sdiff = s2 - s1
    nint = j2 - j1
    r1 = sm(j2 + 1) - sm(j2)
    call solve(nint, sdiff, r1, rat)
    xd = r1
    do j = j1 + 1, j2 - 1
      jsub = j2 + j1 - j
      snew(jsub) = snew(jsub + 1) - xd
      xd = xd*rat
    end do
    if (j1 /= 1) snew(j1) = snew(j1) + 0.01*(snew(j1 + 1) - snew(j1))
    do j = j1, j2
      call intp(ngap, sdist, xint, snew(j), ansx(j))
      call intp(ngap, sdist, rint, snew(j), ansr(j))
    end do
    do j = j1, j2
      XSURF(j, k) = ansx(j)
      RSURF(j, k) = ansr(j)
    end do
    do j = j1, j2 + nextrap
      th_upp(j) = RT_UPP(j, k)/RSURF(j, k)
      th_thick(j) = RT_THICK(j, k)/RSURF(j, k)
    end do
    if (ifangles == 0) then
      slopet = (th_upp(j2) - 0.5*th_thick(j2) - th_upp(j2 + nextrap) &
                + 0.5*th_thick(j2 + nextrap))/(sm(j2) - sm(j2 + nextrap))
    else
      slopet = tan(beta_up*DEGRAD)/RSURF(j2, k)
    end if
    rtheta_le = RT_UPP(j2, k) - 0.5*RT_THICK(j2, k)
    do j = j1, j2 - 1
      RT_UPP(j, k) = rtheta_le + RSURF(j, k)*slopet*(snew(j) - snew(j2))
    end do
  end function grid_up

  ! GRID_DOWN, mf:9201-9431: J1..J2 downstream of a trailing edge at J1, a
  ! geometric progression from the last blade spacing on the gap curve; the
  ! blade's own cusp is kept and the zero-thickness wake line runs on at the
  ! camber-line slope of the last NEXTRAP points (IFCUSP = 0), or a new cusp
  ! of LCUSP points replaces the trailing edge LCUSPUP points upstream of it
  ! (IFCUSP = 1; ICUSP = 0 centred, 1 or -1 keeping the I = 1 or I = IM
  ! surface straight).
  logical function grid_down(k, j1, j2, s1, s2, ngap, sdist, xint, rint, sm, nextrap, beta_dwn1, beta_dwn2, &
                             ifcusp, icusp_r, lcusp_r, lcuspup_r, ifangles) result(bad)
    integer, intent(in) :: k, j1, j2, ngap, nextrap, ifcusp, icusp_r, lcusp_r, lcuspup_r, ifangles
    real, intent(in) :: s1, s2, sdist(:), xint(:), rint(:), sm(:), beta_dwn1, beta_dwn2
    integer, parameter :: LUP = 20
    real :: snew(0:JM), ansx(JM), ansr(JM), th_upp(JM), th_thick(JM), th_mid(JM)
    real :: sdiff, r1, rat, xd, rd, dsmer, slopes, slopetk, slopep, slopecent, betadown1, betadown2
    real :: fdown, betadown, dsdist, rthik, tkgrad, ygrad, ythickte, ymidte, ymid, clength, plength, ytnew
    integer :: j, js, jstrt, jcs, nint, jcuspend
    bad = .false.
    if (j2 - j1 < 1 .or. j1 - nextrap < 1 .or. &
        (ifcusp /= 0 .and. (j1 - LUP - 1 < 1 .or. lcuspup_r > LUP - 2 .or. lcuspup_r < 0 .or. lcusp_r < 1 &
                            .or. j1 - lcuspup_r + lcusp_r > j2))) then
      bad = fail('GRID_DOWN: the row or its gap is too short for NEXTRAP_TE or the cusp'); return
    end if
    snew(j1) = s1
    snew(j2) = s2
    sdiff = s2 - s1
    nint = j2 - j1
    r1 = sm(j1) - sm(j1 - 1)
    call solve(nint, sdiff, r1, rat)
    xd = r1
    do j = j1 + 1, j2 - 1
      snew(j) = snew(j - 1) + xd
      xd = xd*rat
    end do
    do j = j1, j2
      call intp(ngap, sdist, xint, snew(j), ansx(j))
      call intp(ngap, sdist, rint, snew(j), ansr(j))
    end do
    do j = j1, j2
      XSURF(j, k) = ansx(j)
      RSURF(j, k) = ansr(j)
    end do
    if (ifcusp == 0) then
      if (ifangles == 0) then
        ! the arc length restarts from zero one point before J1-NEXTRAP;
        ! with input angles the progression's SNEW stays
        jstrt = j1 - nextrap
        snew(jstrt - 1) = 0.0
        do j = jstrt, j2
          xd = XSURF(j, k) - XSURF(j - 1, k)
          rd = RSURF(j, k) - RSURF(j - 1, k)
          snew(j) = snew(j - 1) + sqrt(xd*xd + rd*rd)
          th_upp(j) = RT_UPP(j, k)/RSURF(j, k)
          th_thick(j) = RT_THICK(j, k)/RSURF(j, k)
          th_mid(j) = th_upp(j) - 0.5*th_thick(j)
        end do
        dsmer = snew(j1) - snew(jstrt)
        slopes = (th_upp(j1) - th_upp(jstrt))/dsmer
        slopetk = (th_thick(j1) - th_thick(jstrt))/dsmer
        slopep = slopes - slopetk
        slopecent = 0.5*(slopes + slopep)
        betadown1 = slopecent
        betadown2 = betadown1
      else
        betadown1 = tan(beta_dwn1*DEGRAD)/RSURF(j1, k)
        betadown2 = tan(beta_dwn2*DEGRAD)/RSURF(j2, k)
      end if
      ! the end of the deck's own cusp: the first point of zero thickness
      js = j1
      do j = j1, j2
        if (RT_THICK(j, k) < 1.0E-4*sdiff) then
          js = j
          exit
        end if
      end do
      do j = js + 1, j2
        fdown = (snew(j) - snew(j1))/(snew(j2) - snew(j1))
        betadown = betadown1 + fdown*(betadown2 - betadown1)
        dsdist = snew(j) - snew(j - 1)
        RT_UPP(j, k) = RT_UPP(j - 1, k) + RSURF(j, k)*betadown*dsdist
        RT_THICK(j, k) = 0.0
      end do
      return
    end if
    ! a new cusp
    jstrt = j1 - LUP
    snew(jstrt - 1) = 0.0
    do j = jstrt, j2
      xd = XSURF(j, k) - XSURF(j - 1, k)
      rd = RSURF(j, k) - RSURF(j - 1, k)
      snew(j) = snew(j - 1) + sqrt(xd*xd + rd*rd)
      th_upp(j) = RT_UPP(j, k)/RSURF(j, k)
      th_thick(j) = RT_THICK(j, k)/RSURF(j, k)
      th_mid(j) = th_upp(j) - 0.5*th_thick(j)
    end do
    js = jstrt + 1
    if (js > j1) js = j1
    ! the start of the trailing-edge thinning
    jcs = j1
    do j = js, j1
      rthik = th_thick(j)/th_thick(j - 1)
      if (rthik < 0.9) then
        jcs = j - 1
        exit
      end if
    end do
    if (jcs - 2 < jstrt) then
      bad = fail('GRID_DOWN: the trailing-edge thinning starts too close to J1-20'); return
    end if
    ! the blade run on straight (camber line and thickness) from JCS to J1
    tkgrad = (th_thick(jcs) - th_thick(jcs - 2))/(snew(jcs) - snew(jcs - 2))
    ygrad = (th_mid(jcs) - th_mid(jcs - 2))/(snew(jcs) - snew(jcs - 2))
    do j = jcs, j1
      th_mid(j) = th_mid(jcs) + ygrad*(snew(j) - snew(jcs))
      th_thick(j) = th_thick(jcs) + tkgrad*(snew(j) - snew(jcs))
      if (th_thick(j) < 0.0) th_thick(j) = 0.0
      th_upp(j) = th_mid(j) + 0.5*th_thick(j)
      RT_UPP(j, k) = RSURF(j, k)*th_upp(j)
      RT_THICK(j, k) = RSURF(j, k)*th_thick(j)
    end do
    ! slopes at the new trailing edge, LCUSPUP points before J1
    jstrt = j1 - lcuspup_r
    dsmer = snew(jstrt) - snew(jstrt - 2)
    slopes = (th_upp(jstrt) - th_upp(jstrt - 2))/dsmer
    slopetk = (th_thick(jstrt) - th_thick(jstrt - 2))/dsmer
    slopep = slopes - slopetk
    slopecent = 0.5*(slopes + slopep)
    if (ifangles == 0) then
      betadown1 = slopecent
      betadown2 = slopecent
    else
      betadown1 = tan(beta_dwn1*DEGRAD)/RSURF(jstrt, k)
      betadown2 = tan(beta_dwn2*DEGRAD)/RSURF(jstrt, k)
    end if
    ythickte = RT_THICK(jstrt, k)
    if (ythickte < 0.0) ythickte = 0.0
    ymidte = RT_UPP(jstrt, k) - 0.5*ythickte
    do j = jstrt + 1, j2
      fdown = (snew(j) - snew(jstrt))/(snew(j2) - snew(jstrt))
      betadown = betadown1 + fdown*(betadown2 - betadown1)
      dsdist = snew(j) - snew(j - 1)
      ymid = ymidte + RSURF(j, k)*slopecent*(snew(j) - snew(jstrt))
      clength = snew(jstrt + lcusp_r) - snew(jstrt)
      plength = snew(j) - snew(jstrt)
      if (plength > clength) plength = clength
      jcuspend = jstrt + lcusp_r
      if ((jcuspend - j) >= 0) then
        ytnew = ythickte*(1. - plength/clength)
        RT_THICK(j, k) = ytnew
        if (icusp_r == 0) RT_UPP(j, k) = ymid + 0.5*ytnew
        if (icusp_r == 1) RT_UPP(j, k) = RT_UPP(j - 1, k) + RSURF(j, k)*slopes*dsdist
        if (icusp_r == -1) RT_UPP(j, k) = RT_UPP(j - 1, k) - RT_THICK(j - 1, k) + RSURF(j, k)*slopep*dsdist + ytnew
      else
        RT_UPP(j, k) = RT_UPP(j - 1, k) + RSURF(j, k)*betadown*dsdist
        RT_THICK(j, k) = 0.0
      end if
    end do
  end function grid_down

  ! INPINT, mf:9073-9195: inlet conditions from the KIN input points onto KM
  ! stream surfaces.
  subroutine inpint()
    real :: sumfn(KM), sumf_in(KIN), ans(KM)
    integer :: k
    sumfn(1) = 0.0
    do k = 2, KM
      sumfn(k) = sumfn(k - 1) + FR(k - 1)
    end do
    do k = 2, KM
      sumfn(k) = sumfn(k)/sumfn(KM)
    end do
    sumf_in(1) = 0.0
    do k = 2, KIN
      sumf_in(k) = sumf_in(k - 1) + FR_IN(k - 1)
    end do
    do k = 2, KIN
      sumf_in(k) = sumf_in(k)/sumf_in(KIN)
    end do
    do k = 1, KM
      call intp(KIN, sumf_in, PO1(1:KIN), sumfn(k), ans(k))
    end do
    call grow(PO1); PO1(1:KM) = ans
    do k = 1, KM
      call intp(KIN, sumf_in, TO1(1:KIN), sumfn(k), ans(k))
    end do
    call grow(TO1); TO1(1:KM) = ans
    do k = 1, KM
      call intp(KIN, sumf_in, VTIN(1:KIN), sumfn(k), ans(k))
    end do
    call grow(VTIN); VTIN(1:KM) = ans
    do k = 1, KM
      call intp(KIN, sumf_in, VM1(1:KIN), sumfn(k), ans(k))
    end do
    call grow(VM1); VM1(1:KM) = ans
    if (IPOUT == 3) then
      do k = 1, KM
        call intp(KIN, sumf_in, PD(1:KIN), sumfn(k), ans(k))
      end do
      call grow(PD); PD(1:KM) = ans
    end if
    do k = 1, KM
      call intp(KIN, sumf_in, BR(1:KIN), sumfn(k), ans(k))
    end do
    call grow(BR); BR(1:KM) = ans
    do k = 1, KM
      call intp(KIN, sumf_in, BS(1:KIN), sumfn(k), ans(k))
    end do
    call grow(BS); BS(1:KM) = ans
  contains
    subroutine grow(a)
      real, allocatable, intent(inout) :: a(:)
      real, allocatable :: b(:)
      if (size(a) >= KM) return
      allocate (b(KM)); b = 0.; b(1:size(a)) = a
      call move_alloc(b, a)
    end subroutine grow
  end subroutine inpint

  ! ---------------------------------------------------------------------------
  ! NEW_INTPOL, mf:7493-7610: KM stream surfaces from the NSECS_IN sections,
  ! with the tip-gap spacing of a row squeezed into KTIPS..KTIPE.
  subroutine new_intpol()
    real :: xint(NSECS_IN), yint(NSECS_IN), rtup_int(NSECS_IN), rttk_int(NSECS_IN), frmod(KM), qodist(NSECS_IN)
    real :: sm(JM, NSECS_IN), xd, rd, frac, xarg, qospan
    double precision :: sumfr, tipgap, tipspace, fac
    integer :: k, j, nr, j1, jl, jt, je, ktip, ncell
    do k = 1, NSECS_IN
      sm(1, k) = 0.0
      do j = 2, JM
        xd = XSURF(j, k) - XSURF(j - 1, k)
        rd = RSURF(j, k) - RSURF(j - 1, k)
        sm(j, k) = sm(j - 1, k) + sqrt(xd*xd + rd*rd)
      end do
    end do
    tipgap = 0.d0
    do j = 1, JM
      nr = NROW(j)
      j1 = JSTART(nr); jl = JLE(nr); jt = JTE(nr); je = JMIX(nr)
      do k = 1, KMM1
        frmod(k) = FR(k)
      end do
      frmod(KM) = 0.0
      if (KTIPS(nr) > 0) then
        ktip = NSECS_IN
        if (KTIPS(nr) == 1) ktip = 1
        fac = 1.0
        if (j >= j1 .and. j < jl) then
          fac = (sm(j, ktip) - sm(j1, ktip))/(sm(jl, ktip) - sm(j1, ktip))
          tipgap = FRACTIP1(nr)
        end if
        if (j >= jl .and. j <= jt) then
          fac = 1.0
          frac = (sm(j, ktip) - sm(jl, ktip))/(sm(jt, ktip) - sm(jl, ktip))
          tipgap = FRACTIP1(nr) + frac*(FRACTIP2(nr) - FRACTIP1(nr))
        end if
        if (j > jt .and. j <= je) then
          fac = (sm(je, ktip) - sm(j, ktip))/(sm(je, ktip) - sm(jt, ktip))
          tipgap = FRACTIP2(nr)
        end if
        sumfr = 0.0
        ncell = KTIPE(nr) - KTIPS(nr)
        tipspace = tipgap/ncell
        do k = KTIPS(nr), KTIPE(nr) - 1
          sumfr = sumfr + FR(k)
        end do
        do k = KTIPS(nr), KTIPE(nr) - 1
          frmod(k) = tipspace*fac + (1. - fac)*FR(k)
        end do
        do k = KS1(nr), KS2(nr)
          frmod(k) = FR(k)*((1. - tipgap)/(1. - sumfr)*fac + (1. - fac))
        end do
      end if
      qodist(1) = 0.0
      do k = 1, NSECS_IN
        xint(k) = XSURF(j, k)
        yint(k) = RSURF(j, k)
        rtup_int(k) = RT_UPP(j, k)
        rttk_int(k) = RT_THICK(j, k)
        if (k > 1) then
          rd = RSURF(j, k) - RSURF(j, k - 1)
          xd = XSURF(j, k) - XSURF(j, k - 1)
          qodist(k) = qodist(k - 1) + sqrt(xd*xd + rd*rd)
        end if
      end do
      xarg = 0.0
      qospan = qodist(NSECS_IN)
      do k = 1, KM
        call intp(NSECS_IN, qodist, xint, xarg, X(j, k))
        call intp(NSECS_IN, qodist, yint, xarg, R(j, k))
        call intp(NSECS_IN, qodist, rtup_int, xarg, RT_UPP(j, k))
        call intp(NSECS_IN, qodist, rttk_int, xarg, RT_THICK(j, k))
        xarg = xarg + frmod(k)*qospan
      end do
    end do
  end subroutine new_intpol

  ! STEPUP, mf:7916-8002 (ideal gas)
  subroutine stepup(relax)
    real, intent(in) :: relax
    real :: tlim, vlim, relax1, vsq, hstat, tstatic, gamnow, wtrel, wsq, v_sonic, vplusc
    real :: amachp, amachl, avmach, stepnew
    integer :: i, j, k, ip1, kp1, nr, jj
    tlim = 0.1*TO1(KMID)
    vlim = 0.01*VM1(KMID)*VM1(KMID)
    relax1 = 1.0 - relax
    do k = 1, KM
      do j = 1, JM
        do i = 1, IM
          vsq = VX(i, j, k)*VX(i, j, k) + VT(i, j, k)*VT(i, j, k) + VR(i, j, k)*VR(i, j, k)
          hstat = HO(i, j, k) - 0.5*vsq
          tstatic = hstat*RCP
          gamnow = GA
          if (tstatic < tlim) tstatic = tlim
          wtrel = WT(i, j, k)
          wsq = vsq - (VT(i, j, k)*VT(i, j, k) - wtrel*wtrel)
          if (wsq < vlim) wsq = vlim
          v_sonic = sqrt(gamnow*RGAS*tstatic)
          vplusc = sqrt(wsq) + v_sonic
          TEMP1(i, j, k) = VSOUND/vplusc
        end do
      end do
    end do
    do i = 1, IMM1
      ip1 = i + 1
      do k = 1, KMM1
        kp1 = k + 1
        amachp = TEMP1(i, 1, k) + TEMP1(ip1, 1, k) + TEMP1(i, 1, kp1) + TEMP1(ip1, 1, kp1)
        do j = 2, JM
          amachl = TEMP1(i, j, k) + TEMP1(i, j, kp1) + TEMP1(ip1, j, k) + TEMP1(ip1, j, kp1)
          avmach = 0.125*(amachl + amachp)
          amachp = amachl
          stepnew = BSTEP(i, j, k)*avmach
          STEP(i, j, k) = relax*stepnew + relax1*STEP(i, j, k)
          RSTEP(i, j, k) = STEP(i, j, k)/BSTEP(i, j, k)
        end do
      end do
    end do
    do nr = 1, NRWSM1
      jj = JMIX(nr)
      do k = 1, KMM1
        do i = 1, IMM1
          STEP(i, jj + 1, k) = STEP(i, jj + 2, k)
          RSTEP(i, jj + 1, k) = 0.0
        end do
      end do
    end do
  end subroutine stepup

  ! ---------------------------------------------------------------------------
  ! The restart file in place of the initial guess, mf:6919-6999 (ITIMST < 5,
  ! IFGAS = 0): the conserved variables and the SA variable from the
  ! plotting/restart file 'flow_out' of the current directory (the record
  ! layout OUTPUT writes, mf:5951-5976; its step count and ROSUB are not
  ! used), the secondary variables from them, and, for IPOUT <= 0, the exit
  ! pressure profile shifted to the file's pitch-averaged exit pressures.
  logical function read_restart() result(bad)
    integer :: u, ios, i, j, k, nstep_file
    real :: eke, tstatic, pdiff
    real, allocatable :: rosub(:,:,:), pdown(:)
    bad = .false.
    allocate (rosub(IM, JM, KM), pdown(KM))
    open (newunit=u, file='flow_out', form='unformatted', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      bad = fail('IF_RESTART set but no restart file flow_out'); return
    end if
    read (u, iostat=ios) nstep_file
    if (ios == 0) read (u, iostat=ios) RO
    if (ios == 0) read (u, iostat=ios) ROVX
    if (ios == 0) read (u, iostat=ios) ROVR
    if (ios == 0) read (u, iostat=ios) ROVT
    if (ios == 0) read (u, iostat=ios) ROE
    if (ios == 0) read (u, iostat=ios) rosub
    if (ios == 0) read (u, iostat=ios) TRANS_DYN_VIS
    close (u)
    if (ios /= 0) then
      bad = fail('restart file flow_out: short or of another grid'); return
    end if
    do k = 1, KM
      do j = 1, JM
        do i = 1, IM
          VX(i, j, k) = ROVX(i, j, k)/RO(i, j, k)
          VR(i, j, k) = ROVR(i, j, k)/RO(i, j, k)
          VT(i, j, k) = ROVT(i, j, k)/RO(i, j, k)
          RORVT(i, j, k) = ROVT(i, j, k)*R(j, k)
          eke = 0.5*(VX(i, j, k)*VX(i, j, k) + VR(i, j, k)*VR(i, j, k) + VT(i, j, k)*VT(i, j, k))
          tstatic = (ROE(i, j, k)/RO(i, j, k) - eke)/CV
          T_STATIC(i, j, k) = tstatic
          HO(i, j, k) = CP*tstatic + eke
          P(i, j, k) = RO(i, j, k)*RGAS*tstatic
          PEFF(i, j, k) = P(i, j, k)
        end do
      end do
      pdown(k) = 0.0
      do i = 1, IMM1
        pdown(k) = pdown(k) + FP(i)*0.5*(P(i, JM, k) + P(i + 1, JM, k))
      end do
    end do
    if (IPOUT <= 0) then
      pdiff = 0.0
      if (IPOUT == 0) pdiff = PD(1) - pdown(1)
      if (IPOUT == -1) pdiff = PD(KM) - pdown(KM)
      do k = 1, KM
        PD(k) = pdown(k) + pdiff
      end do
    end if
  end function read_restart

  ! SETUP, mf:6280-7487 (up to the viscous-model initialisation)
  logical function setup() result(bad)
    integer :: i, j, k, nr, jledge, jtedge, kavg, jstt, mark, jp1, nsb, jsb, i1, j1, k1, i2, j2, k2
    integer :: istart, iend, jstrt, jendd, kstart, kend
    real :: sum, shift, rtmid, thick, pitch, gap, schord, tchord, dsup, dsdwn, fac, x1, x2, r1, r2, t1, t2
    real :: rarea, vinsq, ps, prat, ts, ros, facj, area, ain, vs, vtgrid, eke, smin, attot, perpt, perpq
    real :: astot, perps, volb, perpmin, perpavg, dr1, dx1, drm, dxm, avgspan
    integer, allocatable :: jstart1(:), jend1(:), jstart2(:), jend2(:)
    bad = .false.
    NRWSM1 = NROWS - 1
    IMID = ifix(0.5*IM)
    KMID = ifix(0.5*KM)
    allocate (BSSIN(KM), BSCOS(KM), BRSIN(KM), BRCOS(KM), RPO1(KM))
    do k = 1, KM
      BSSIN(k) = sin(BS(k)*DEGRAD)
      BSCOS(k) = cos(BS(k)*DEGRAD)
      BRSIN(k) = sin(BR(k)*DEGRAD)
      BRCOS(k) = cos(BR(k)*DEGRAD)
    end do
    sum = 0.0
    do k = 1, KMM1
      sum = sum + FR(k)
    end do
    do k = 1, KMM1
      FR(k) = FR(k)/sum
    end do
    FSPAN(1) = 0.5*FR(1)
    do k = 2, KMM1
      FSPAN(k) = FSPAN(k - 1) + 0.5*(FR(k) + FR(k - 1))
    end do
    sum = 0.0
    do i = 1, IMM1
      sum = sum + FP(i)
    end do
    do i = 1, IMM1
      FP(i) = FP(i)/sum
    end do
    FU = 0.; FD = 0.; FKU = 0.; FKD = 0.
    do i = 2, IMM1
      FU(i) = FP(i)/(FP(i) + FP(i - 1))
      FD(i) = FP(i - 1)/(FP(i) + FP(i - 1))
    end do
    FU(1) = FP(1)/FP(2)
    FU(IM) = FP(IMM1)/FP(IMM2)
    do k = 2, KMM1
      FKU(k) = FR(k)/(FR(k) + FR(k - 1))
      FKD(k) = FR(k - 1)/(FR(k) + FR(k - 1))
    end do
    FKU(1) = FR(1)/FR(2)
    FKU(KM) = FR(KMM1)/FR(KMM2)
    ! corner distribution weights, mf:6382-6410
    allocate (FTL(IM, KM), FBL(IM, KM), FTR(IM, KM), FBR(IM, KM), FACUP(JM), FACDWN(JM))
    FTL = 0.; FBL = 0.; FTR = 0.; FBR = 0.
    do k = 1, KMM1
      do i = 1, IMM1
        FTL(i, k) = 0.125; FBL(i, k) = 0.125; FTR(i, k) = 0.125; FBR(i, k) = 0.125
      end do
    end do
    do k = 1, KMM1
      FTL(1, k) = 0.25; FBL(1, k) = 0.25; FTR(IMM1, k) = 0.25; FBR(IMM1, k) = 0.25
    end do
    do i = 1, IMM1
      FBL(i, 1) = 0.25; FBR(i, 1) = 0.25; FTL(i, KMM1) = 0.25; FTR(i, KMM1) = 0.25
    end do
    FTL(1, KMM1) = 0.5; FBL(1, 1) = 0.5; FTR(IMM1, KMM1) = 0.5; FBR(IMM1, 1) = 0.5
    FACUP = 0.; FACDWN = 0.
    do j = 2, JM
      FACUP(j) = 1.0
      FACDWN(j) = 1.0
    end do
    FACUP(2) = 2.0
    FACDWN(JM) = 2.0
    ! grid
    allocate (X(JM, KM), R(JM, KM))
    block
      real, allocatable :: rtu(:,:), rtt(:,:)
      allocate (rtu(JM, KM), rtt(JM, KM))
      call move_alloc(RT_UPP, rtu)      ! sections, [J, NSECS_IN]
      call move_alloc(RT_THICK, rtt)
      allocate (RT_UPP(JM, KM), RT_THICK(JM, KM))
      RT_UPP = 0.; RT_THICK = 0.
      RT_UPP(1:JM, 1:NSECS_IN) = rtu(1:JM, 1:NSECS_IN)
      RT_THICK(1:JM, 1:NSECS_IN) = rtt(1:JM, 1:NSECS_IN)
    end block
    call new_intpol()
    ! align the rows circumferentially at I=1 on mid-span, mf:6450-6456
    shift = 0.0
    do j = 3, JM
      if (INDMIX(j - 1) == 1) shift = (RT_UPP(j - 1, KMID) - RT_UPP(j, KMID))/R(j, KMID)
      do k = 1, KM
        RT_UPP(j, k) = RT_UPP(j, k) + shift*R(j, k)
      end do
    end do
    ! pinched-tip thickness scaling, mf:6460-6470
    do j = 1, JM
      nr = NROW(j)
      do k = 1, KM
        rtmid = RT_UPP(j, k) - 0.5*RT_THICK(j, k)
        thick = FTHICK(nr, k)*RT_THICK(j, k)
        RT_UPP(j, k) = rtmid + 0.5*thick
        RT_THICK(j, k) = thick
      end do
    end do
    GA1 = GA - 1.0
    FGA = GA1/GA
    RFGA = 1.0/FGA
    RCP = 1.0/CP
    CV = CP/GA
    RCV = 1.0/CV
    RGAS = CP - CV
    SFEX1 = 1. - SFEXIT
    SFEXH = 0.5*SFEXIT
    ! blade-to-blade coordinates, mf:6488-6503
    allocate (THETA(IM, JM, KM), RTHETA(IM, JM, KM), RT_PITCH(JM, KM))
    do j = 1, JM
      pitch = 2.*PI/NBLADE(j)
      do k = 1, KM
        gap = pitch*R(j, k) - RT_THICK(j, k)
        RTHETA(1, j, k) = RT_UPP(j, k)
        THETA(1, j, k) = RT_UPP(j, k)/R(j, k)
        RT_PITCH(j, k) = pitch*R(j, k)
        do i = 2, IM
          RTHETA(i, j, k) = RTHETA(i - 1, j, k) + FP(i - 1)*gap
          THETA(i, j, k) = RTHETA(i, j, k)/R(j, k)
        end do
      end do
    end do
    allocate (SMERID(JM, KM), DX(JM, KM), DR(JM, KM), DS(JM, KM))
    DSMIN = 1.0E06
    do k = 1, KM
      SMERID(1, k) = 0.0
      do j = 2, JM
        DR(j, k) = (R(j, k) - R(j - 1, k))
        DX(j, k) = (X(j, k) - X(j - 1, k))
        DS(j, k) = sqrt(DR(j, k)*DR(j, k) + DX(j, k)*DX(j, k))
        SMERID(j, k) = SMERID(j - 1, k) + DS(j, k)
        if (DS(j, k) < DSMIN) DSMIN = DS(j, k)
      end do
      DX(1, k) = DX(2, k)
      DS(1, k) = DS(2, k)
      DR(1, k) = DR(2, k)
    end do
    do nr = 1, NROWS
      jledge = JLE(nr)
      jtedge = JTE(nr)
      schord = SMERID(jtedge, KMID) - SMERID(jledge, KMID)
      tchord = RTHETA(1, jledge, KMID) - RTHETA(1, jtedge, KMID)
      CHORD(nr) = sqrt(schord*schord + tchord*tchord)
    end do
    allocate (RAVG_CELL(JM, KM), UBLADE(JM, KM), FMUP(JM, KM), FMDN(JM, KM))
    RAVG_CELL = 0.; FMUP = 0.; FMDN = 0.
    do j = 2, JM
      do k = 1, KMM1
        RAVG_CELL(j, k) = 0.25*(R(j, k) + R(j - 1, k) + R(j - 1, k + 1) + R(j, k + 1))
      end do
    end do
    do j = 1, JM
      do k = 1, KM
        UBLADE(j, k) = WRAD(j)*R(j, k)
      end do
    end do
    do k = 1, KM
      do j = 2, JMM1
        dsup = SMERID(j, k) - SMERID(j - 1, k)
        dsdwn = SMERID(j + 1, k) - SMERID(j, k)
        FMDN(j, k) = dsdwn/(dsup + dsdwn)
        FMUP(j, k) = dsup/(dsup + dsdwn)
      end do
    end do
    ! theta extrapolated onto the mixing plane, mf:6570-6580
    allocate (TEMP1(IM, JM, KM), TEMP2(IM, JM, KM))
    TEMP1 = 0.; TEMP2 = 0.
    do j = 1, JM
      if (INDMIX(j) /= 1) cycle
      do k = 1, KM
        fac = DS(j + 1, k)/DS(j + 2, k)
        do i = 1, IM
          TEMP1(i, j, k) = THETA(i, j + 1, k) + fac*(THETA(i, j + 1, k) - THETA(i, j + 2, k))
          TEMP2(i, j, k) = R(j, k)*TEMP1(i, j, k)
        end do
      end do
    end do
    ! face areas and volumes, mf:6609-6728
    allocate (AQX(IM, JM, KM), AQR(IM, JM, KM), AQTOT(IM, JM, KM), ABX(IM, JM, KM), ABR(IM, JM, KM))
    allocate (ABTOT(IM, JM, KM), ASX(IM, JM, KM), ASR(IM, JM, KM), VOL(IM, JM, KM), VOLOR(IM, JM, KM))
    allocate (ABT(JM, KM), WRABT(JM, KM))
    AQX = 0.; AQR = 0.; AQTOT = 0.; ABX = 0.; ABR = 0.; ABTOT = 0.; ASX = 0.; ASR = 0.; VOL = 0.
    VOLOR = 0.; ABT = 0.; WRABT = 0.
    do j = 1, JM
      do i = 1, IMM1
        do k = 1, KMM1
          x1 = X(j, k + 1) - X(j, k)
          x2 = x1
          r1 = R(j, k + 1) - R(j, k)
          r2 = r1
          t1 = R(j, k + 1)*(THETA(i + 1, j, k + 1) - THETA(i, j, k))
          t2 = R(j, k + 1)*(THETA(i, j, k + 1) - THETA(i, j, k)) - R(j, k)*(THETA(i + 1, j, k) - THETA(i, j, k))
          AQX(i, j, k) = -0.5*(r2*t1 - r1*t2)
          AQR(i, j, k) = -0.5*(x1*t2 - x2*t1)
          AQTOT(i, j, k) = sqrt(AQX(i, j, k)*AQX(i, j, k) + AQR(i, j, k)*AQR(i, j, k))
        end do
        AQX(i, j, KM) = AQX(i, j, KMM1)
        AQR(i, j, KM) = AQR(i, j, KMM1)
      end do
    end do
    do j = 1, JM
      do k = 1, KM
        AQX(IM, j, k) = AQX(IMM1, j, k)
        AQR(IM, j, k) = AQR(IMM1, j, k)
      end do
    end do
    do j = 2, JM
      do k = 1, KMM1
        x1 = X(j - 1, k + 1) - X(j, k)
        x2 = X(j, k + 1) - X(j - 1, k)
        r1 = R(j - 1, k + 1) - R(j, k)
        r2 = R(j, k + 1) - R(j - 1, k)
        ABT(j, k) = -0.5*(x1*r2 - x2*r1)
        WRABT(j, k) = WRAD(j)*0.25*(R(j, k) + R(j - 1, k) + R(j - 1, k + 1) + R(j, k + 1))*ABT(j, k)
        do i = 1, IM
          t1 = R(j - 1, k + 1)*(THETA(i, j - 1, k + 1) - THETA(i, j, k))
          t2 = R(j, k + 1)*(THETA(i, j, k + 1) - THETA(i, j, k)) - R(j - 1, k)*(THETA(i, j - 1, k) - THETA(i, j, k))
          if (INDMIX(j - 1) == 1) then
            t1 = R(j - 1, k + 1)*(TEMP1(i, j - 1, k + 1) - THETA(i, j, k))
            t2 = R(j, k + 1)*(THETA(i, j, k + 1) - THETA(i, j, k)) - R(j - 1, k)*(TEMP1(i, j - 1, k) - THETA(i, j, k))
          end if
          ABX(i, j, k) = -0.5*(r1*t2 - r2*t1)
          ABR(i, j, k) = -0.5*(x2*t1 - x1*t2)
          ABTOT(i, j, k) = sqrt(ABX(i, j, k)**2 + ABR(i, j, k)**2 + ABT(j, k)**2)
          if (i == IM) cycle
          VOL(i, j, k) = ABT(j, k)*0.25*(RTHETA(i + 1, j, k) + RTHETA(i + 1, j - 1, k) + RTHETA(i + 1, j - 1, k + 1) &
                         + RTHETA(i + 1, j, k + 1) - RTHETA(i, j, k) - RTHETA(i, j - 1, k) - RTHETA(i, j - 1, k + 1) &
                         - RTHETA(i, j, k + 1))
          if (INDMIX(j - 1) == 1) &
            VOL(i, j, k) = ABT(j, k)*0.25*(RTHETA(i + 1, j, k) + TEMP2(i + 1, j - 1, k) + TEMP2(i + 1, j - 1, k + 1) &
                           + RTHETA(i + 1, j, k + 1) - RTHETA(i, j, k) - TEMP2(i, j - 1, k) - TEMP2(i, j - 1, k + 1) &
                           - RTHETA(i, j, k + 1))
        end do
      end do
    end do
    do j = 2, JM
      do i = 1, IMM1
        do k = 1, KM
          x1 = X(j - 1, k) - X(j, k)
          x2 = x1
          r1 = R(j - 1, k) - R(j, k)
          r2 = r1
          t1 = R(j - 1, k)*(THETA(i, j - 1, k) - THETA(i, j, k)) - R(j, k)*(THETA(i + 1, j, k) - THETA(i, j, k))
          t2 = R(j - 1, k)*(THETA(i + 1, j - 1, k) - THETA(i, j, k))
          if (INDMIX(j - 1) == 1) then
            t1 = R(j - 1, k)*(TEMP1(i, j - 1, k) - THETA(i, j, k)) - R(j, k)*(THETA(i + 1, j, k) - THETA(i, j, k))
            t2 = R(j - 1, k)*(TEMP1(i + 1, j - 1, k) - THETA(i, j, k))
          end if
          ASX(i, j, k) = 0.5*(r1*t2 - r2*t1)
          ASR(i, j, k) = 0.5*(x2*t1 - x1*t2)
        end do
      end do
    end do
    do k = 1, KMM1
      do j = 2, JM
        do i = 1, IMM1
          if (VOL(i, j, k) < 0.0) then
            bad = fail('negative cell volumes in the grid'); return
          end if
          rarea = ABR(i, j, k) - ABR(i + 1, j, k) + AQR(i, j, k) - AQR(i, j - 1, k) + ASR(i, j, k) - ASR(i, j, k + 1)
          VOLOR(i, j, k) = -0.125*rarea
        end do
      end do
    end do
    ! initial guess, mf:6760-6913
    if (IPOUT /= 3) then
      PD(1) = PDOWN_HUB
      do k = 2, KM
        PD(k) = PD(k - 1) + FR(k - 1)*(PDOWN_TIP - PDOWN_HUB)
      end do
    end if
    allocate (RO(IM, JM, KM), P(IM, JM, KM), T_STATIC(IM, JM, KM), ROAVG_CELL(IM, JM, KM), VX(IM, JM, KM))
    allocate (VR(IM, JM, KM), VT(IM, JM, KM), WT(IM, JM, KM), HO(IM, JM, KM))
    VX = 0.; VR = 0.; VT = 0.; WT = 0.; HO = 0.
    do k = 1, KM
      RPO1(k) = 1./PO1(k)
      VR(1, 1, k) = VM1(k)*(R(2, k) - R(1, k))/DS(1, k)
      vinsq = VM1(k)*VM1(k) + VTIN(k)*VTIN(k)
      do j = 1, JM
        ps = PGUESS(j)
        prat = ps/PO1(k)
        ts = (prat**FGA)*TO1(k)
        ros = ps/ts/RGAS
        do i = 1, IM
          P(i, j, k) = ps
          RO(i, j, k) = ros
          ROAVG_CELL(i, j, k) = ros
          T_STATIC(i, j, k) = ts
        end do
      end do
    end do
    kavg = int(0.75*KM)
    do k = 1, KM
      do i = 1, IM
        VT(i, 1, k) = VTIN(k)
        WT(i, 1, k) = VTIN(k) - UBLADE(1, k)
        HO(i, 1, k) = CP*TO1(k)
        facj = 0.9
        do j = 2, JM
          if (j > 2) then
            if (INDTE(j - 2) == 1) facj = 1.0
          end if
          area = sqrt(AQX(IMID, j, kavg)*AQX(IMID, j, kavg) + AQR(IMID, j, kavg)*AQR(IMID, j, kavg))*NBLADE(j)
          if (j == 2) ain = area
          vs = VM1(kavg)*ain*RO(IMID, 1, kavg)/(area*RO(IMID, j, kavg))
          vtgrid = vs*R(j, k)*(THETA(i, j, k) - THETA(i, j - 1, k))/DS(j, k) + UBLADE(j, k)
          VT(i, j, k) = (1. - facj)*vtgrid + facj*VT(i, j - 1, k)*R(j - 1, k)/R(j, k)
          if (INDMIX(j - 1) == 1) VT(i, j, k) = VT(i, j - 1, k)
          if (j < JLE(1)) VT(i, j, k) = VTIN(k)*R(1, k)/R(j, k)
          WT(i, j, k) = VT(i, j, k) - UBLADE(j, k)
          VX(i, j, k) = vs*DX(j, k)/DS(j, k)
          VR(i, j, k) = vs*DR(j, k)/DS(j, k)
          HO(i, j, k) = HO(i, j - 1, k) + WRAD(j)*(R(j, k)*VT(i, j, k) - R(j - 1, k)*VT(i, j - 1, k))
          if (INDLE(j) == 1) facj = 0.9
        end do
        VR(i, 1, k) = VR(i, 2, k)
        VX(i, 1, k) = VX(i, 2, k)
        WT(i, 1, k) = WT(i, 2, k)
      end do
    end do
    PREFF = 0.5*(P(IMID, 1, KMID) + P(IMID, JM, KMID))
    allocate (ROE(IM, JM, KM), ROVX(IM, JM, KM), ROVR(IM, JM, KM), ROVT(IM, JM, KM), ROWT(IM, JM, KM))
    allocate (RORVT(IM, JM, KM), PEFF(IM, JM, KM), DRO(IM, JM, KM), DROE(IM, JM, KM), DROVX(IM, JM, KM))
    allocate (DROVR(IM, JM, KM), DRORVT(IM, JM, KM))
    do j = 1, JM
      do k = 1, KM
        do i = 1, IM
          eke = 0.5*(VX(i, j, k)**2 + VT(i, j, k)**2 + VR(i, j, k)**2)
          ROE(i, j, k) = P(i, j, k)/(GA - 1) + eke*RO(i, j, k)
          ros = RO(i, j, k)
          ROVR(i, j, k) = ros*VR(i, j, k)
          ROVT(i, j, k) = ros*VT(i, j, k)
          ROWT(i, j, k) = ros*WT(i, j, k)
          RORVT(i, j, k) = ROVT(i, j, k)*R(j, k)
          ROVX(i, j, k) = ros*VX(i, j, k)
          PEFF(i, j, k) = P(i, j, k) - PREFF
          DRO(i, j, k) = 0.0; DROE(i, j, k) = 0.0; DROVX(i, j, k) = 0.0; DROVR(i, j, k) = 0.0
          DRORVT(i, j, k) = 0.0
        end do
      end do
    end do
    allocate (TRANS_DYN_VIS(IM, JM, KM))
    TRANS_DYN_VIS = 0.0
    if (IF_RESTART /= 0) then
      bad = read_restart()
      if (bad) return
    end if
    ! time steps, mf:7012-7058
    VSOUND = sqrt(RGAS*GA*TO1(KMID))
    allocate (STEP(IM, JM, KM), BSTEP(IM, JM, KM), RSTEP(IM, JM, KM))
    STEP = 0.; BSTEP = 0.; RSTEP = 0.
    do j = 2, JM
      do k = 1, KMM1
        do i = 1, IMM1
          smin = DSMIN
          attot = sqrt(ABT(j, k)*ABT(j, k) + ABX(i, j, k)*ABX(i, j, k) + ABR(i, j, k)*ABR(i, j, k))
          perpt = VOL(i, j, k)/attot
          perpq = VOL(i, j, k)/AQTOT(i, j, k)
          astot = sqrt(ASX(i, j, k)*ASX(i, j, k) + ASR(i, j, k)*ASR(i, j, k))
          perps = VOL(i, j, k)/astot
          smin = amin1(perpt, perpq, perps)
          STEP(i, j, k) = smin*CFL/VOL(i, j, k)/VSOUND
          BSTEP(i, j, k) = STEP(i, j, k)
          RSTEP(i, j, k) = 1.0
        end do
      end do
    end do
    do nr = 1, NRWSM1
      j = JMIX(nr)
      do k = 1, KMM1
        do i = 1, IMM1
          STEP(i, j + 1, k) = 0.5*(STEP(i, j, k) + STEP(i, j + 2, k))
          BSTEP(i, j + 1, k) = STEP(i, j + 1, k)
          RSTEP(i, j + 1, k) = 0.0
        end do
      end do
    end do
    call stepup(1.0)
    ! multigrid block maps, mf:7067-7180
    allocate (KB1(KM), KB2(KM), IB1(IM), IB2(IM), JB1(JM), JB2(JM), JSBLK(JM))
    KB1 = 0; KB2 = 0; IB1 = 0; IB2 = 0
    do k = 1, KMM1
      KB2(k) = 1 + (k - 1)/KRBB
      KB1(k) = 1 + (k - 1)/KR
    end do
    NKB1 = KB1(KMM1)
    NKB2 = KB2(KMM1)
    do i = 1, IMM1
      IB2(i) = 1 + (i - 1)/IRBB
      IB1(i) = 1 + (i - 1)/IR
    end do
    NIB1 = IB1(IMM1)
    NIB2 = IB2(IMM1)
    jstt = 2
    JB1(1) = 1
    JB1(2) = 1
    do j = 3, JM
      mark = 0
      if (INDMIX(j) == 1 .or. j == JM) mark = 1
      if ((mod((j - jstt), JR) == 0) .and. (mark == 0)) then
        JB1(j) = JB1(j - 1) + 1
      else
        JB1(j) = JB1(j - 1)
      end if
      if (INDMIX(j - 1) == 1) JB1(j) = 0
      if (INDMIX(j - 2) == 1) then
        JB1(j) = JB1(j - 2) + 1
        jstt = j
      end if
    end do
    jstt = 2
    JB2(1) = 1
    JB2(2) = 1
    do j = 3, JM
      jp1 = j + 1
      if (j == JM) jp1 = JM
      mark = 0
      if (INDMIX(j) == 1 .or. j == JM) mark = 1
      if (INDMIX(jp1) == 1 .or. (j + 1) == JM) mark = 1
      if ((mod((j - jstt), JRBB) == 0) .and. (mark == 0)) then
        JB2(j) = JB2(j - 1) + 1
      else
        JB2(j) = JB2(j - 1)
      end if
      if (INDMIX(j - 1) == 1) JB2(j) = 0
      if (INDMIX(j - 2) == 1) then
        JB2(j) = JB2(j - 2) + 1
        jstt = j
      end if
    end do
    NJB1 = JB1(JM)
    NJB2 = JB2(JM)
    allocate (jstart1(NJB1 + 1), jend1(NJB1 + 1), jstart2(NJB2 + 1), jend2(NJB2 + 1))
    jstart1(1) = 2
    jstart2(1) = 2
    do j = 2, JM
      if (JB1(j) == 0) JB1(j) = NJB1 + 1
      if (JB2(j) == 0) JB2(j) = NJB2 + 1
      if (JB1(j) /= JB1(j - 1)) jstart1(JB1(j)) = j
      jend1(JB1(j)) = j
      if (JB2(j) /= JB2(j - 1)) jstart2(JB2(j)) = j
      jend2(JB2(j)) = j
    end do
    jstart1(NJB1 + 1) = JM; jend1(NJB1 + 1) = JM
    jstart2(NJB2 + 1) = JM; jend2(NJB2 + 1) = JM
    nsb = 1
    do j = 1, JM
      JSBLK(j) = nsb
      if (INDLE(j) == 1) nsb = nsb + 1
      if (INDMID(j) == 1) nsb = nsb + 1
      if (INDTE(j) == 1) nsb = nsb + 1
    end do
    NSBLK = JSBLK(JM)
    ! multigrid block time steps, mf:7186-7279
    allocate (STEP1(NIB1, NJB1 + 1, NKB1), STEP2(NIB2, NJB2 + 1, NKB2), STEPSBK(NSBLK))
    STEP1 = 0.; STEP2 = 0.; STEPSBK = 0.
    do i1 = 1, NIB1
      istart = (i1 - 1)*IR + 1
      iend = istart + IR - 1
      if (iend > IMM1) iend = IMM1
      do j1 = 1, NJB1
        jstrt = jstart1(j1)
        jendd = jend1(j1)
        do k1 = 1, NKB1
          kstart = (k1 - 1)*KR + 1
          kend = kstart + KR - 1
          if (kend > KMM1) kend = KMM1
          perpt = 0.0; perpq = 0.0; perps = 0.0; volb = 0.0
          do i = istart, iend
            do k = kstart, kend
              do j = jstrt, jendd
                attot = sqrt(ABT(j, k)*ABT(j, k) + ABX(i, j, k)*ABX(i, j, k) + ABR(i, j, k)*ABR(i, j, k))
                perpt = perpt + VOL(i, j, k)/attot
                perpq = perpq + VOL(i, j, k)/AQTOT(i, j, k)
                astot = sqrt(ASX(i, j, k)*ASX(i, j, k) + ASR(i, j, k)*ASR(i, j, k))
                perps = perps + VOL(i, j, k)/astot
                volb = volb + VOL(i, j, k)
              end do
            end do
          end do
          perps = perps/(IR*JR)
          perpt = perpt/(JR*KR)
          perpq = perpq/(IR*KR)
          perpmin = amin1(perps, perpt, perpq)
          STEP1(i1, j1, k1) = CFL*FBLK1*perpmin/VSOUND/volb
          STEP1(i1, NJB1 + 1, k1) = 0.0
        end do
      end do
    end do
    do i2 = 1, NIB2
      istart = (i2 - 1)*IRBB + 1
      iend = istart + IRBB - 1
      if (iend > IMM1) iend = IMM1
      do j2 = 1, NJB2
        jstrt = jstart2(j2)
        jendd = jend2(j2)
        do k2 = 1, NKB2
          kstart = (k2 - 1)*KRBB + 1
          kend = kstart + KRBB - 1
          if (kend > KMM1) kend = KMM1
          perpt = 0.0; perpq = 0.0; perps = 0.0; volb = 0.0
          do i = istart, iend
            do k = kstart, kend
              do j = jstrt, jendd
                attot = sqrt(ABT(j, k)*ABT(j, k) + ABX(i, j, k)*ABX(i, j, k) + ABR(i, j, k)*ABR(i, j, k))
                perpt = perpt + VOL(i, j, k)/attot
                perpq = perpq + VOL(i, j, k)/AQTOT(i, j, k)
                astot = sqrt(ASX(i, j, k)*ASX(i, j, k) + ASR(i, j, k)*ASR(i, j, k))
                perps = perps + VOL(i, j, k)/astot
                volb = volb + VOL(i, j, k)
              end do
            end do
          end do
          perps = perps/(IRBB*JRBB)
          perpt = perpt/(JRBB*KRBB)
          perpq = perpq/(IRBB*KRBB)
          perpmin = amin1(perps, perpt, perpq)
          STEP2(i2, j2, k2) = CFL*FBLK2*perpmin/VSOUND/volb
          STEP2(i2, NJB2 + 1, k2) = 0.0
        end do
      end do
    end do
    perpq = 0.0
    volb = 0.0
    do j = 2, JM
      jsb = JSBLK(j)
      do i = 1, IMM1
        do k = 1, KMM1
          perpq = perpq + VOL(i, j, k)/AQTOT(i, j, k)
          volb = volb + VOL(i, j, k)
        end do
      end do
      if (j == JM) then
        perpavg = perpq/(IMM1*KMM1)
        STEPSBK(jsb) = CFL*FBLK3*perpavg/VSOUND/volb
        cycle
      end if
      if (JSBLK(j + 1) /= jsb) then
        perpavg = perpq/(IMM1*KMM1)
        STEPSBK(jsb) = CFL*FBLK3*perpavg/VSOUND/volb
        perpq = 0.0
        volb = 0.0
      end if
    end do
    allocate (XFORCE(IM, JM, KM), RFORCE(IM, JM, KM), TFORCE(IM, JM, KM), QSOURCE(IM, JM, KM))
    XFORCE = 0.0; RFORCE = 0.0; TFORCE = 0.0; QSOURCE = 0.0
    allocate (FLOWX(IM, JM, KM), FLOWT(IM, JM, KM), FLOWR(IM, JM, KM))
    FLOWX = 0.0; FLOWT = 0.0; FLOWR = 0.0
    dr1 = R(1, KM) - R(1, 1)
    dx1 = X(1, KM) - X(1, 1)
    drm = R(JM, KM) - R(JM, 1)
    dxm = X(JM, KM) - X(JM, 1)
    avgspan = 0.5*(sqrt(dx1*dx1 + dr1*dr1) + sqrt(dxm*dxm + drm*drm))
    VORT_MAX = 5000.*VM1(KMID)/avgspan
    ! wall distances and mixing lengths; the viscous models use them
    ! (mf:7340), an inviscid run does not; with host_xlength off the
    ! caller forms them on the device (mh_set_xlength) and they stay zero here
    if (ILOS /= 0 .and. host_xlength) then
      call set_xlength()
    else
      allocate (DWALLSQ(IM, JM, KM), XLENGTH(IM, JM, KM), FST_RAT(IM, JM, KM), XLIMIT(JM))
      DWALLSQ = 0.0; XLENGTH = 0.0; FST_RAT = 0.0; XLIMIT = 0.0
    end if
  end function setup

  !> Distance of every node to the nearest end wall and blade surface, the
  !> mixing lengths derived from it and the squared cell wall distance,
  !> restating SET_XLENGTH (mf:12194-12517): the nearest wall point is sought
  !> over J +/- 20 (and K +/- 5 on the blades), and the distance is taken
  !> normal to the wall there.
  subroutine set_xlength()
    integer, parameter :: jrange = 20, krange = 5
    real, allocatable :: dist_min(:,:,:), xl(:,:,:), t3(:,:,:)
    integer :: i, j, k, nr, iwall, j1, j2, k1, k2, jwall, jsurf, ksurf, jmin, kmin, knorm, if_found, iside, isurf
    real :: pitch, dist_ref, hubdist, tipdist, dmin, xdif, rdif, distsq, atot, xnorm, rnorm, tnorm, tdif, tdifsq
    real :: endwall_dist, blade_dist, sdist(2), distsqrt, xllim, dblend, expon, fblend, x_length, ratio
    real :: avg_xlength, avg_fstrat, avgdist
    allocate (dist_min(IM, JM, KM), xl(IM, JM, KM), t3(IM, JM, KM), XLIMIT(JM))
    allocate (DWALLSQ(IM, JM, KM), XLENGTH(IM, JM, KM), FST_RAT(IM, JM, KM))
    DWALLSQ = 0.0; XLENGTH = 0.0; FST_RAT = 0.0; XLIMIT = 0.0
    xllim = 0.0; jmin = 1; kmin = 1
    do k = 1, KM
      do j = 1, JM
        nr = NROW(j)
        pitch = 6.283185*R(j, k)/NBLADE(j)
        dist_ref = 0.5*pitch
        do i = 1, IM
          ! end walls: nearest hub / casing point over J +/- jrange, then the
          ! distance normal to the wall there (mf:12227-12287)
          iwall = IMID
          j1 = max(j - jrange, 2)
          j2 = min(j + jrange, JM)
          dmin = 1.0E10
          do jwall = j1, j2
            xdif = X(j, k) - X(jwall, 1)
            rdif = R(j, k) - R(jwall, 1)
            distsq = xdif*xdif + rdif*rdif
            if (distsq < dmin) then
              dmin = distsq
              jmin = jwall
            end if
          end do
          atot = sqrt(ASX(iwall, jmin, 1)*ASX(iwall, jmin, 1) + ASR(iwall, jmin, 1)*ASR(iwall, jmin, 1))
          xnorm = ASX(iwall, jmin, 1)/atot
          rnorm = ASR(iwall, jmin, 1)/atot
          xdif = X(j, k) - X(jmin, 1)
          rdif = R(j, k) - R(jmin, 1)
          hubdist = abs(xdif*xnorm + rdif*rnorm)
          if (IBOUND == 1 .or. IBOUND > 2) hubdist = dist_ref
          dmin = 1.0E10
          do jwall = j1, j2
            xdif = X(j, k) - X(jwall, KM)
            rdif = R(j, k) - R(jwall, KM)
            distsq = xdif*xdif + rdif*rdif
            if (distsq < dmin) then
              dmin = distsq
              jmin = jwall
            end if
          end do
          atot = sqrt(ASX(iwall, jmin, KM)*ASX(iwall, jmin, KM) + ASR(iwall, jmin, KM)*ASR(iwall, jmin, KM))
          xnorm = ASX(iwall, jmin, KM)/atot
          rnorm = ASR(iwall, jmin, KM)/atot
          xdif = X(j, k) - X(jmin, KM)
          rdif = R(j, k) - R(jmin, KM)
          tipdist = abs(xdif*xnorm + rdif*rnorm)
          if (IBOUND >= 2) tipdist = dist_ref
          endwall_dist = hubdist*tipdist/(hubdist + tipdist)
          ! blade surfaces I = 1 and I = IM of this row (mf:12296-12402)
          j1 = j - jrange
          if (j1 < 2) j1 = 2
          if (j1 < JSTART(nr)) j1 = JSTART(nr)
          if (j1 >= JTE(nr)) j1 = JTE(nr) - 1
          j2 = j + jrange
          if (j2 > JM) j2 = JM
          if (j2 > JMIX(nr)) j2 = JMIX(nr)
          if (j2 <= JLE(nr)) j2 = JLE(nr) + 1
          k1 = max(k - krange, 1)
          k2 = min(k + krange, KM)
          do iside = 1, 2
            isurf = 1
            if (iside == 2) isurf = IM
            dmin = 1.0E10
            if_found = 0
            do jsurf = j1, j2
              do ksurf = k1, k2
                xdif = X(j, k) - X(jsurf, ksurf)
                rdif = R(j, k) - R(jsurf, ksurf)
                tdifsq = 4.*R(j, k)*R(jsurf, ksurf)*sin((THETA(i, j, k) - THETA(isurf, jsurf, ksurf))/2.)**2
                distsq = xdif*xdif + rdif*rdif + tdifsq
                if (distsq < dmin) then
                  dmin = distsq
                  jmin = jsurf
                  kmin = ksurf
                  if_found = 1
                end if
              end do
            end do
            if (if_found == 1) then
              knorm = kmin
              if (kmin == KM) knorm = KM - 1
              atot = sqrt(ABX(isurf, jmin, knorm)*ABX(isurf, jmin, knorm) + ABR(isurf, jmin, knorm)* &
                          ABR(isurf, jmin, knorm) + ABT(jmin, knorm)*ABT(jmin, knorm))
              xnorm = ABX(isurf, jmin, knorm)/atot
              rnorm = ABR(isurf, jmin, knorm)/atot
              tnorm = ABT(jmin, knorm)/atot
              xdif = X(j, k) - X(jmin, kmin)
              rdif = R(j, k) - R(jmin, kmin)
              tdif = 2.*sqrt(R(j, k)*R(jmin, kmin))*sin((THETA(i, j, k) - THETA(isurf, jmin, kmin))/2.)
              sdist(iside) = abs(xdif*xnorm + rdif*rnorm + tdif*tnorm)
              if (jmin <= JLE(nr) .or. jmin >= JTE(nr)) sdist(iside) = sqrt(xdif*xdif + rdif*rdif + tdif*tdif)
            else
              sdist(iside) = dist_ref
            end if
          end do
          blade_dist = sdist(1)*sdist(2)/(sdist(1) + sdist(2))
          ! combined distance (mf:12408-12424)
          if (k > 1 .and. k < KM) then
            distsqrt = sqrt(endwall_dist*endwall_dist + blade_dist*blade_dist)
            if (distsqrt < 1.0E-10) distsqrt = 1.0E-10
            dist_min(i, j, k) = blade_dist*endwall_dist/distsqrt
          else
            dist_min(i, j, k) = 0.0
          end if
          if (KM == 2) dist_min(i, j, k) = blade_dist
          if (IM == 2) dist_min(i, j, k) = endwall_dist
          ! mixing-length limit varying along the row (mf:12428-12449)
          if (i == IMID .and. k == KMID) then
            if (j <= JLE(nr)) xllim = XLLIM_IN(nr) + (SMERID(j, k) - SMERID(JSTART(nr), k)) &
                                      /(SMERID(JLE(nr), k) - SMERID(JSTART(nr), k))*(XLLIM_LE(nr) - XLLIM_IN(nr))
            if (j > JLE(nr) .and. j <= JTE(nr)) xllim = XLLIM_LE(nr) + (SMERID(j, k) - SMERID(JLE(nr), k)) &
                                      /(SMERID(JTE(nr), k) - SMERID(JLE(nr), k))*(XLLIM_TE(nr) - XLLIM_LE(nr))
            if (j > JTE(nr)) xllim = XLLIM_TE(nr) + (SMERID(j, k) - SMERID(JTE(nr), k)) &
                                      /(SMERID(JMIX(nr), k) - SMERID(JTE(nr), k))*(XLLIM_DN(nr) - XLLIM_TE(nr))
            XLIMIT(j) = 2.0*dist_min(i, j, k)*xllim
          end if
        end do
      end do
    end do
    ! blended mixing length and the free-stream damping function (mf:12460-12477)
    dblend = 0.6666666
    expon = dblend/(1.0 - dblend)
    fblend = (1.0 - dblend)*dblend**expon
    do j = 1, JM
      do k = 1, KM
        do i = 1, IM
          x_length = dist_min(i, j, k)
          ratio = x_length/XLIMIT(j)
          if (ratio > dblend) x_length = XLIMIT(j)*(1.0 - fblend*ratio**(-expon))
          xl(i, j, k) = x_length
          t3(i, j, k) = (x_length/XLIMIT(j))**4
        end do
      end do
    end do
    ! cell averages (mf:12482-12508)
    do k = 1, KMM1
      do j = 2, JM
        nr = NROW(j)
        do i = 1, IMM1
          avg_xlength = 0.125*(xl(i, j, k) + xl(i, j, k + 1) + xl(i, j - 1, k) + xl(i, j - 1, k + 1) + xl(i + 1, j, k) &
                               + xl(i + 1, j, k + 1) + xl(i + 1, j - 1, k) + xl(i + 1, j - 1, k + 1))
          XLENGTH(i, j, k) = 0.41*0.41*avg_xlength*avg_xlength
          if (j == JSTART(nr)) XLENGTH(i, j, k) = 0.0
          avg_fstrat = 0.125*(t3(i, j, k) + t3(i, j, k + 1) + t3(i, j - 1, k) + t3(i, j - 1, k + 1) + t3(i + 1, j, k) &
                              + t3(i + 1, j, k + 1) + t3(i + 1, j - 1, k) + t3(i + 1, j - 1, k + 1))
          FST_RAT(i, j, k) = avg_fstrat*FSTURB(nr)
          if (j == JSTART(nr)) FST_RAT(i, j, k) = 0.0
          avgdist = 0.125*(dist_min(i, j, k) + dist_min(i, j, k + 1) + dist_min(i, j - 1, k) + dist_min(i, j - 1, k + 1) &
                           + dist_min(i + 1, j, k) + dist_min(i + 1, j, k + 1) + dist_min(i + 1, j - 1, k) &
                           + dist_min(i + 1, j - 1, k + 1))
          DWALLSQ(i, j, k) = avgdist*avgdist
        end do
      end do
    end do
    deallocate (dist_min, xl, t3)
  end subroutine set_xlength

  subroutine reset()
    ! release everything from a previous load
    if (allocated(FP)) deallocate (FP, FR, FU, FD, FKU, FKD, FSPAN)
    if (allocated(NSTAGE)) deallocate (NSTAGE)
    if (allocated(JSTART)) deallocate (JSTART, JMIX, JLE, JTE, JLED, KTIPS, KTIPE, KS1, KS2, IF_CUSP, IF_ANGLES, &
                                       FRACTIP1, FRACTIP2, FTHICK, CHORD, JTRAN_I1, JTRAN_IM, JTRAN_K1, JTRAN_KM, &
                                       ICUSP, LCUSP, LCUSPUP, BETAUP, BETADWN1, BETADWN2)
    if (allocated(NUP_I1)) deallocate (NUP_I1, NUP_IM, N_WAKE, SEP_THIK, SEP_DRAG)
    if (allocated(XSURF)) deallocate (XSURF, RSURF)
    if (allocated(RT_UPP)) deallocate (RT_UPP, RT_THICK)
    if (allocated(IND)) deallocate (IND, INDLE, INDTE, INDMIX, INDMID, NBLADE, NROW, WRAD, WHUB, WTIP, PGUESS)
    if (allocated(FR_IN)) deallocate (FR_IN, PO1, TO1, VTIN, VM1, BS, BR, PD)
    if (allocated(XLLIM_I1)) deallocate (XLLIM_I1, XLLIM_IM, XLLIM_K1, XLLIM_KM, XLLIM_DWN, XLLIM_UP, XLLIM_IN, &
                                         XLLIM_LE, XLLIM_TE, XLLIM_DN, FSTURB, TURBVIS_DAMP, ROUGH_H, ROUGH_T, &
                                         ROUGH_L, ROUGH_U)
    if (allocated(BSSIN)) deallocate (BSSIN, BSCOS, BRSIN, BRCOS, RPO1)
    if (allocated(FTL)) deallocate (FTL, FBL, FTR, FBR, FACUP, FACDWN)
    if (allocated(X)) deallocate (X, R)
    if (allocated(THETA)) deallocate (THETA, RTHETA, RT_PITCH)
    if (allocated(SMERID)) deallocate (SMERID, DX, DR, DS)
    if (allocated(RAVG_CELL)) deallocate (RAVG_CELL, UBLADE, FMUP, FMDN)
    if (allocated(TEMP1)) deallocate (TEMP1, TEMP2)
    if (allocated(AQX)) deallocate (AQX, AQR, AQTOT, ABX, ABR, ABTOT, ASX, ASR, VOL, VOLOR, ABT, WRABT)
    if (allocated(RO)) deallocate (RO, P, T_STATIC, ROAVG_CELL, VX, VR, VT, WT, HO)
    if (allocated(ROE)) deallocate (ROE, ROVX, ROVR, ROVT, ROWT, RORVT, PEFF, DRO, DROE, DROVX, DROVR, DRORVT)
    if (allocated(STEP)) deallocate (STEP, BSTEP, RSTEP)
    if (allocated(KB1)) deallocate (KB1, KB2, IB1, IB2, JB1, JB2, JSBLK)
    if (allocated(STEP1)) deallocate (STEP1, STEP2, STEPSBK)
    if (allocated(XFORCE)) deallocate (XFORCE, RFORCE, TFORCE, QSOURCE, FLOWX, FLOWT, FLOWR)
    if (allocated(DWALLSQ)) deallocate (DWALLSQ, XLENGTH, FST_RAT, XLIMIT)
    if (allocated(TRANS_DYN_VIS)) deallocate (TRANS_DYN_VIS)
    ! the mixing-plane parameters are read only for IFMIX /= 0 (mf:476-489)
    ! and the y+ damping range only for ILOS /= 0 (mf:3642-3649); otherwise
    ! they keep the zero of a fresh COMMON block, whatever a deck loaded
    ! before set
    RFMIX = 0.0; FEXTRAP = 0.0; FSMTHB = 0.0; FANGLE = 0.0; YPLAM = 0.0; YPTURB = 0.0
    ! the Spalart-Allmaras factors are read only for ILOS >= 200; otherwise
    ! they keep the zero of a fresh COMMON block (multall-open-18.3.f:3338)
    FAC_STMIX = 0.0; FAC_ST0 = 0.0; FAC_ST1 = 0.0; FAC_ST2 = 0.0; FAC_ST3 = 0.0; FAC_SFVIS = 0.0
    FAC_VORT = 0.0; FAC_PGRAD = 0.0
  end subroutine reset

  !> Read a deck from an open unit and build the case.  Returns .true. on
  !> failure (message in mh_host_err).
  !> The plotting grid file SETUP writes (mf:6585-6599): NSTEPS_MAX, the
  !> grid size, CP and GA, the leading/trailing-edge marker, WRAD and NBLADE
  !> per J, then X, R and the I line of R*THETA for every (J, K).
  logical function write_grid_out(path) result(bad)
    character(len=*), intent(in) :: path
    integer :: u, ios, j, k, i
    integer, allocatable :: indlete(:)
    bad = .false.
    open (newunit=u, file=path, form='unformatted', status='replace', iostat=ios)
    if (ios /= 0) then
      bad = fail('cannot open ' // trim(path)); return
    end if
    allocate (indlete(JM))
    do j = 1, JM
      indlete(j) = 0
      if (INDLE(j) == 1) indlete(j) = 1
      if (INDTE(j) == 1) indlete(j) = 1
    end do
    write (u) NSTEPS_MAX
    write (u) IM, JM, KM
    write (u) CP, GA
    write (u) (indlete(j), j = 1, JM)
    write (u) (WRAD(j), j = 1, JM)
    write (u) (NBLADE(j), j = 1, JM)
    do j = 1, JM
      do k = 1, KM
        write (u) X(j, k), R(j, k), (RTHETA(i, j, k), i = 1, IM)
      end do
    end do
    close (u)
  end function write_grid_out

  logical function mh_host_load_unit(u) result(bad)
    integer, intent(in) :: u
    call reset()
    mh_host_err = ''
    bad = read_deck(u)
    if (bad) return
    bad = setup()
  end function mh_host_load_unit

  logical function mh_host_load_path(path) result(bad)
    character(len=*), intent(in) :: path
    integer :: u, ios
    open (newunit=u, file=path, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      mh_host_err = 'cannot open deck '//trim(path)
      bad = .true.
      return
    end if
    bad = mh_host_load_unit(u)
    close (u)
  end function mh_host_load_path

end module mh_host
