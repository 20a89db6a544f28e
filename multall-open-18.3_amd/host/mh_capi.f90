!> C-callable view of the host set-up (mh_host): load a deck and copy the
!> resulting arrays and scalars out by their reference names.  Used by the
!> Python front end; the Fortran driver uses mh_host directly.
module mh_capi
  use iso_c_binding
  use mh_host
  implicit none
contains

  function cstr(p) result(s)
    character(kind=c_char), intent(in) :: p(*)
    character(len=:), allocatable :: s
    integer :: n
    n = 0
    do while (p(n + 1) /= c_null_char)
      n = n + 1
    end do
    allocate (character(len=n) :: s)
    s = transfer(p(1:n), s)
  end function cstr

  !> Load and set up a deck.  Returns 0 or -1 (see mhh_error).
  integer(c_int) function mhh_load(path) bind(C, name='mhh_load')
    character(kind=c_char), intent(in) :: path(*)
    mhh_load = 0
    if (mh_host_load_path(cstr(path))) mhh_load = -1
  end function mhh_load

  !> Write the loaded deck's plotting grid file (grid_out) to path.
  integer(c_int) function mhh_write_grid_out(path) bind(C, name='mhh_write_grid_out')
    character(kind=c_char), intent(in) :: path(*)
    mhh_write_grid_out = 0
    if (write_grid_out(cstr(path))) mhh_write_grid_out = -1
  end function mhh_write_grid_out

  !> 1: the next loads form DWALLSQ/XLENGTH/FST_RAT on the host (SET_XLENGTH,
  !> the default); 0: they leave them zero for the library's mh_set_xlength.
  subroutine mhh_set_host_xlength(flag) bind(C, name='mhh_set_host_xlength')
    integer(c_int), value :: flag
    host_xlength = flag /= 0
  end subroutine mhh_set_host_xlength

  !> Copy the last error message (NUL-terminated, truncated to n bytes).
  subroutine mhh_error(buf, n) bind(C, name='mhh_error')
    character(kind=c_char), intent(out) :: buf(*)
    integer(c_int), value :: n
    integer :: i, m
    m = min(int(n) - 1, len_trim(mh_host_err))
    do i = 1, m
      buf(i) = mh_host_err(i:i)
    end do
    buf(max(m, 0) + 1) = c_null_char
  end subroutine mhh_error

  !> Scalar by reference name; returns 0, or -1 if unknown.
  integer(c_int) function mhh_scalar(name, v) bind(C, name='mhh_scalar')
    character(kind=c_char), intent(in) :: name(*)
    real(c_double), intent(out) :: v
    mhh_scalar = 0
    select case (cstr(name))
    case ('IM'); v = IM
    case ('JM'); v = JM
    case ('KM'); v = KM
    case ('NROWS'); v = NROWS
    case ('IMID'); v = IMID
    case ('KMID'); v = KMID
    case ('NIB1'); v = NIB1
    case ('NJB1'); v = NJB1
    case ('NKB1'); v = NKB1
    case ('NIB2'); v = NIB2
    case ('NJB2'); v = NJB2
    case ('NKB2'); v = NKB2
    case ('NSBLK'); v = NSBLK
    case ('ITIMST'); v = ITIMST
    case ('NRSMTH'); v = NRSMTH
    case ('IF_RESTART'); v = IF_RESTART
    case ('NSTEPS_MAX'); v = NSTEPS_MAX
    case ('NCHANGE'); v = NCHANGE
    case ('IR'); v = IR
    case ('JR'); v = JR
    case ('KR'); v = KR
    case ('IRBB'); v = IRBB
    case ('JRBB'); v = JRBB
    case ('KRBB'); v = KRBB
    case ('IFMIX'); v = IFMIX
    case ('IN_PRESS'); v = IN_PRESS
    case ('IN_VTAN'); v = IN_VTAN
    case ('IN_VR'); v = IN_VR
    case ('IN_FLOW'); v = IN_FLOW
    case ('IF_REPEAT'); v = IF_REPEAT
    case ('IPOUT'); v = IPOUT
    case ('NSFEXIT'); v = NSFEXIT
    case ('ILOS'); v = ILOS
    case ('NLOS'); v = NLOS
    case ('IBOUND'); v = IBOUND
    case ('NMIXUP'); v = NMIXUP
    case ('IFGAS'); v = IFGAS
    case ('CP'); v = CP
    case ('GA'); v = GA
    case ('F1'); v = F1
    case ('F2'); v = F2
    case ('F3'); v = F3
    case ('RSMTH'); v = RSMTH
    case ('CFL'); v = CFL
    case ('DAMPIN'); v = DAMPIN
    case ('MACHLIM'); v = MACHLIM
    case ('F_PDOWN'); v = F_PDOWN
    case ('CONLIM'); v = CONLIM
    case ('SFXIN'); v = SFXIN
    case ('SFTIN'); v = SFTIN
    case ('FAC_4TH'); v = FAC_4TH
    case ('FBLK1'); v = FBLK1
    case ('FBLK2'); v = FBLK2
    case ('FBLK3'); v = FBLK3
    case ('RFMIX'); v = RFMIX
    case ('FEXTRAP'); v = FEXTRAP
    case ('FSMTHB'); v = FSMTHB
    case ('FANGLE'); v = FANGLE
    case ('RFIN'); v = RFIN
    case ('NINMOD'); v = NINMOD
    case ('RFINBC'); v = RFINBC
    case ('PO_IN_MID'); v = PO_IN_MID
    case ('VT_IN_MID'); v = VT_IN_MID
    case ('FLOWIN'); v = FLOWIN
    case ('RFLOW'); v = RFLOW
    case ('SFEXIT'); v = SFEXIT
    case ('REYNO'); v = REYNO
    case ('RF_VIS'); v = RF_VIS
    case ('FTRANS'); v = FTRANS
    case ('TURBVIS_LIM'); v = TURBVIS_LIM
    case ('PRANDTL'); v = PRANDTL
    case ('YPLUSWALL'); v = YPLUSWALL
    case ('PDOWN_HUB'); v = PDOWN_HUB
    case ('PDOWN_TIP'); v = PDOWN_TIP
    case ('FACMIXUP'); v = FACMIXUP
    case ('GA1'); v = GA1
    case ('FGA'); v = FGA
    case ('RCP'); v = RCP
    case ('CV'); v = CV
    case ('RGAS'); v = RGAS
    case ('VSOUND'); v = VSOUND
    case ('VORT_MAX'); v = VORT_MAX
    case ('YPLAM'); v = YPLAM
    case ('YPTURB'); v = YPTURB
    case ('CFWALL'); v = CFWALL
    case ('FAC_STMIX'); v = FAC_STMIX
    case ('FAC_ST0'); v = FAC_ST0
    case ('FAC_ST1'); v = FAC_ST1
    case ('FAC_ST2'); v = FAC_ST2
    case ('FAC_ST3'); v = FAC_ST3
    case ('FAC_SFVIS'); v = FAC_SFVIS
    case ('FAC_VORT'); v = FAC_VORT
    case ('FAC_PGRAD'); v = FAC_PGRAD
    case default
      v = 0; mhh_scalar = -1
    end select
  end function mhh_scalar

  !> Copy a named array (float32 or int32, dense, I fastest) into dest of
  !> capacity n elements.  Returns the element count, or -1 if unknown, or
  !> -2 if n is too small.
  integer(c_long_long) function mhh_array(name, dest, n) bind(C, name='mhh_array')
    character(kind=c_char), intent(in) :: name(*)
    type(c_ptr), value :: dest
    integer(c_long_long), value :: n
    real(c_float), pointer :: f(:)
    integer(c_int), pointer :: ii(:)
    character(len=:), allocatable :: nm
    nm = cstr(name)
    mhh_array = -1
    select case (nm)
    case ('RO'); call putr3(RO)
    case ('ROVX'); call putr3(ROVX)
    case ('ROVR'); call putr3(ROVR)
    case ('RORVT'); call putr3(RORVT)
    case ('ROE'); call putr3(ROE)
    case ('DRO'); call putr3(DRO)
    case ('DROVX'); call putr3(DROVX)
    case ('DROVR'); call putr3(DROVR)
    case ('DRORVT'); call putr3(DRORVT)
    case ('DROE'); call putr3(DROE)
    case ('ROVT'); call putr3(ROVT)
    case ('ROWT'); call putr3(ROWT)
    case ('P'); call putr3(P)
    case ('HO'); call putr3(HO)
    case ('PEFF'); call putr3(PEFF)
    case ('VX'); call putr3(VX)
    case ('VR'); call putr3(VR)
    case ('VT'); call putr3(VT)
    case ('WT'); call putr3(WT)
    case ('T_STATIC'); call putr3(T_STATIC)
    case ('AQX'); call putr3(AQX)
    case ('AQR'); call putr3(AQR)
    case ('ABX'); call putr3(ABX)
    case ('ABR'); call putr3(ABR)
    case ('ASX'); call putr3(ASX)
    case ('ASR'); call putr3(ASR)
    case ('VOL'); call putr3(VOL)
    case ('VOLOR'); call putr3(VOLOR)
    case ('STEP'); call putr3(STEP)
    case ('BSTEP'); call putr3(BSTEP)
    case ('RSTEP'); call putr3(RSTEP)
    case ('XFORCE'); call putr3(XFORCE)
    case ('RFORCE'); call putr3(RFORCE)
    case ('TFORCE'); call putr3(TFORCE)
    case ('QSOURCE'); call putr3(QSOURCE)
    case ('FLOWX'); call putr3(FLOWX)
    case ('FLOWT'); call putr3(FLOWT)
    case ('FLOWR'); call putr3(FLOWR)
    case ('DWALLSQ'); call putr3(DWALLSQ)
    case ('XLENGTH'); call putr3(XLENGTH)
    case ('FST_RAT'); call putr3(FST_RAT)
    case ('WHUB'); call putr1(WHUB)
    case ('WTIP'); call putr1(WTIP)
    case ('XLLIM_I1'); call putr1(XLLIM_I1)
    case ('XLLIM_IM'); call putr1(XLLIM_IM)
    case ('XLLIM_K1'); call putr1(XLLIM_K1)
    case ('XLLIM_KM'); call putr1(XLLIM_KM)
    case ('XLLIM_DWN'); call putr1(XLLIM_DWN)
    case ('XLLIM_UP'); call putr1(XLLIM_UP)
    case ('ROUGH_H'); call putr1(ROUGH_H)
    case ('ROUGH_T'); call putr1(ROUGH_T)
    case ('ROUGH_L'); call putr1(ROUGH_L)
    case ('ROUGH_U'); call putr1(ROUGH_U)
    case ('CHORD'); call putr1(CHORD)
    case ('THETA'); call putr3(THETA)
    case ('XLLIM_IN'); call putr1(XLLIM_IN)
    case ('XLLIM_LE'); call putr1(XLLIM_LE)
    case ('XLLIM_TE'); call putr1(XLLIM_TE)
    case ('XLLIM_DN'); call putr1(XLLIM_DN)
    case ('JTRAN_I1'); call puti(JTRAN_I1)
    case ('JTRAN_IM'); call puti(JTRAN_IM)
    case ('JTRAN_K1'); call puti(JTRAN_K1)
    case ('JTRAN_KM'); call puti(JTRAN_KM)
    case ('FSTURB'); call putr1(FSTURB)
    case ('TURBVIS_DAMP'); call putr1(TURBVIS_DAMP)
    case ('ROAVG_CELL'); call putr3(ROAVG_CELL)
    case ('TRANS_DYN_VIS'); call putr3(TRANS_DYN_VIS)
    case ('RTHETA'); call putr3(RTHETA)
    case ('AQTOT'); call putr3(AQTOT)
    case ('ABTOT'); call putr3(ABTOT)
    case ('R'); call putr2(R)
    case ('X'); call putr2(X)
    case ('RAVG_CELL'); call putr2(RAVG_CELL)
    case ('ABT'); call putr2(ABT)
    case ('WRABT'); call putr2(WRABT)
    case ('UBLADE'); call putr2(UBLADE)
    case ('DX'); call putr2(DX)
    case ('DR'); call putr2(DR)
    case ('DS'); call putr2(DS)
    case ('SMERID'); call putr2(SMERID)
    case ('RT_UPP'); call putr2(RT_UPP)
    case ('FTL'); call putr2(FTL)
    case ('FBL'); call putr2(FBL)
    case ('FTR'); call putr2(FTR)
    case ('FBR'); call putr2(FBR)
    case ('FP'); call putr1(FP(1:IM))
    case ('FU'); call putr1(FU(1:IM))
    case ('FD'); call putr1(FD(1:IM))
    case ('FR'); call putr1(FR(1:KM))
    case ('FKU'); call putr1(FKU(1:KM))
    case ('FKD'); call putr1(FKD(1:KM))
    case ('PO1'); call putr1(PO1(1:KM))
    case ('TO1'); call putr1(TO1(1:KM))
    case ('RPO1'); call putr1(RPO1(1:KM))
    case ('BSSIN'); call putr1(BSSIN(1:KM))
    case ('BSCOS'); call putr1(BSCOS(1:KM))
    case ('BRSIN'); call putr1(BRSIN(1:KM))
    case ('BRCOS'); call putr1(BRCOS(1:KM))
    case ('VTIN'); call putr1(VTIN(1:KM))
    case ('PD'); call putr1(PD(1:KM))
    case ('VM1'); call putr1(VM1(1:KM))
    case ('FACUP'); call putr1(FACUP)
    case ('FACDWN'); call putr1(FACDWN)
    case ('WRAD'); call putr1(WRAD)
    case ('STEP1'); call putr1(reshape(STEP1, [size(STEP1)]))
    case ('STEP2'); call putr1(reshape(STEP2, [size(STEP2)]))
    case ('STEPSBK'); call putr1(STEPSBK)
    case ('IB1'); call puti(IB1)
    case ('IB2'); call puti(IB2)
    case ('KB1'); call puti(KB1)
    case ('KB2'); call puti(KB2)
    case ('NBLADE'); call puti(NBLADE)
    case ('IND'); call puti(IND)
    case ('INDLE'); call puti(INDLE)
    case ('INDMIX'); call puti(INDMIX)
    case ('INDTE'); call puti(INDTE)
    case ('NROW'); call puti(NROW)
    case ('JSBLK'); call puti(JSBLK)
    case ('JB1'); call puti(JB1)
    case ('JB2'); call puti(JB2)
    case ('JSTART'); call puti(JSTART(1:NROWS + 1))
    case ('JMIX'); call puti(JMIX(1:NROWS + 1))
    case ('JLE'); call puti(JLE(1:NROWS + 1))
    case ('JTE'); call puti(JTE(1:NROWS + 1))
    case ('JLED'); call puti(JLED(1:NROWS + 1))
    case ('KTIPS'); call puti(KTIPS(1:NROWS + 1))
    case ('KTIPE'); call puti(KTIPE(1:NROWS + 1))
    case ('KS1'); call puti(KS1(1:NROWS + 1))
    case ('KS2'); call puti(KS2(1:NROWS + 1))
    case ('NOUT'); call puti(NOUT)
    case ('NSTAGE'); call puti(NSTAGE(1:NROWS))
    case ('IF_CUSP'); call puti(IF_CUSP(1:NROWS))
    case ('NUP_I1'); call puti(NUP_I1(1:NROWS))
    case ('NUP_IM'); call puti(NUP_IM(1:NROWS))
    case ('N_WAKE'); call puti(N_WAKE(1:NROWS))
    case ('SEP_THIK'); call putr1(SEP_THIK(1:NROWS))
    case ('SEP_DRAG'); call putr1(SEP_DRAG(1:NROWS))
    case ('RT_THICK'); call putr2(RT_THICK(1:JM, 1:KM))
    case ('BS'); call putr1(BS(1:KM))
    case ('BR'); call putr1(BR(1:KM))
    end select
  contains
    subroutine putr1(a)
      real, intent(in) :: a(:)
      if (size(a) > n) then
        mhh_array = -2; return
      end if
      call c_f_pointer(dest, f, [size(a)])
      f = a
      mhh_array = size(a)
    end subroutine putr1
    subroutine putr2(a)
      real, intent(in) :: a(:, :)
      call putr1(reshape(a, [size(a)]))
    end subroutine putr2
    subroutine putr3(a)
      real, intent(in) :: a(:, :, :)
      call putr1(reshape(a, [size(a)]))
    end subroutine putr3
    subroutine puti(a)
      integer, intent(in) :: a(:)
      if (size(a) > n) then
        mhh_array = -2; return
      end if
      call c_f_pointer(dest, ii, [size(a)])
      ii = a
      mhh_array = size(a)
    end subroutine puti
  end function mhh_array

end module mh_capi
