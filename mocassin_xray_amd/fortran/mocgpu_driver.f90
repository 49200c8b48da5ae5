! mocgpu_driver -- a MOCASSIN run through mocgpu_mod from Fortran: reads a
! model file (mesh, source, grid, atomic, balance and emission tables and the
! starting state, as written by the package's fortran_io.write_model), or
! builds the model in the library from the reference's data/ files and a
! namelist of input-deck keywords (moc_model_build), runs MCIterationDriver
! on the device and writes Te, Ne, ionDen and the last J.
!
! usage: mocgpu_driver <model file> <output file>
!        mocgpu_driver --namelist <namelist file> <output file>
!        mocgpu_driver --namelist <namelist file> --write-model <model file>
!          (builds the model and writes it, with no device call)
!
! The namelist is one group &model with the keywords of the reference's
! input deck that the path reads (defaults in brackets): dataDir, elements
! (atomic numbers), logAbundances (log10 n(X)/n(H)), nstages [7], maxLev [27],
! XHILimit [0.05], nbins [1000], nuMin [1.0], nuMax [100.0], strictMesh [T],
! lgCompton [T], TStellar [1e5], LStar [1e36 erg/s, 10], nPhotons [1e6],
! plFraction [0], plLoKeV [0.01], plHiKeV [10], nx [16], Rout [cm, 1e18],
! Hdensity [100], TeStart [1e4], H0Start [1e-5], precision ['fp64' or
! 'mixed'], maxIterateMC [5], minConvergence [100], seed, lgAutoPackets [F],
! convIncPercent [0], nPhotIncrease [2], maxPhotons [0].
!
! Model and output files are a stream of records: the 8 characters
! "MOCDUMP1", then per record a 32-character name, an int32 kind (1 real(8),
! 2 integer(4), 3 integer(8)), an int32 0, an int64 count and the values;
! a record named "END" closes the file.
module mocdump_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private
  type :: rec_t
     character(len=32) :: name
     integer :: kind = 0
     integer(c_int64_t) :: n = 0
     real(c_double), allocatable :: d(:)
     integer(c_int32_t), allocatable :: i(:)
     integer(c_int64_t), allocatable :: l(:)
  end type rec_t
  type(rec_t), allocatable, target, save :: recs(:)
  integer, save :: nrec = 0
  public :: dump_read, dptr, iptr, lptr, ival, lval, dval, has, wrec_d, wrec_i, wrec_end, &
       wrec_begin
contains

  subroutine dump_read(path)
    character(len=*), intent(in) :: path
    character(len=8) :: magic
    character(len=32) :: name
    integer(c_int32_t) :: kind, pad
    integer(c_int64_t) :: n
    type(rec_t), allocatable :: tmp(:)
    integer :: u, ios
    open(newunit=u, file=path, access="stream", form="unformatted", status="old", &
         action="read", iostat=ios)
    if (ios /= 0) then
       print*, "! mocgpu_driver: can't open model file: ", trim(path)
       stop
    end if
    read(u) magic
    if (magic /= "MOCDUMP1") then
       print*, "! mocgpu_driver: not a model file: ", trim(path)
       stop
    end if
    allocate(recs(64))
    nrec = 0
    do
       read(u) name, kind, pad, n
       if (trim(name) == "END") exit
       if (nrec == size(recs)) then
          allocate(tmp(2 * size(recs)))
          tmp(1:nrec) = recs(1:nrec)
          call move_alloc(tmp, recs)
       end if
       nrec = nrec + 1
       recs(nrec)%name = name
       recs(nrec)%kind = kind
       recs(nrec)%n = n
       select case (kind)
       case (1)
          allocate(recs(nrec)%d(max(n, 1_c_int64_t)))
          if (n > 0) read(u) recs(nrec)%d(1:n)
       case (2)
          allocate(recs(nrec)%i(max(n, 1_c_int64_t)))
          if (n > 0) read(u) recs(nrec)%i(1:n)
       case (3)
          allocate(recs(nrec)%l(max(n, 1_c_int64_t)))
          if (n > 0) read(u) recs(nrec)%l(1:n)
       case default
          print*, "! mocgpu_driver: bad record kind in ", trim(path)
          stop
       end select
    end do
    close(u)
  end subroutine dump_read

  integer function find(name)
    character(len=*), intent(in) :: name
    integer :: k
    find = 0
    do k = 1, nrec
       if (trim(recs(k)%name) == name) then
          find = k
          return
       end if
    end do
  end function find

  logical function has(name)
    character(len=*), intent(in) :: name
    has = find(name) > 0
  end function has

  integer function need(name, kind)
    character(len=*), intent(in) :: name
    integer, intent(in) :: kind
    need = find(name)
    if (need == 0) then
       print*, "! mocgpu_driver: record missing: ", name
       stop
    end if
    if (recs(need)%kind /= kind) then
       print*, "! mocgpu_driver: record of the wrong kind: ", name
       stop
    end if
  end function need

  ! address of a record's values (null for an absent or empty record)
  type(c_ptr) function dptr(name)
    character(len=*), intent(in) :: name
    integer :: k
    dptr = c_null_ptr
    k = find(name)
    if (k > 0) then
       if (recs(k)%n > 0) dptr = c_loc(recs(k)%d(1))
    end if
  end function dptr

  type(c_ptr) function iptr(name)
    character(len=*), intent(in) :: name
    integer :: k
    iptr = c_null_ptr
    k = find(name)
    if (k > 0) then
       if (recs(k)%n > 0) iptr = c_loc(recs(k)%i(1))
    end if
  end function iptr

  type(c_ptr) function lptr(name)
    character(len=*), intent(in) :: name
    integer :: k
    lptr = c_null_ptr
    k = find(name)
    if (k > 0) then
       if (recs(k)%n > 0) lptr = c_loc(recs(k)%l(1))
    end if
  end function lptr

  integer(c_int32_t) function ival(name, i)
    character(len=*), intent(in) :: name
    integer, intent(in), optional :: i
    integer :: k, j
    k = need(name, 2)
    j = 1
    if (present(i)) j = i
    ival = recs(k)%i(j)
  end function ival

  integer(c_int64_t) function lval(name)
    character(len=*), intent(in) :: name
    lval = recs(need(name, 3))%l(1)
  end function lval

  real(c_double) function dval(name)
    character(len=*), intent(in) :: name
    dval = recs(need(name, 1))%d(1)
  end function dval

  subroutine wrec_begin(u)
    integer, intent(in) :: u
    write(u) "MOCDUMP1"
  end subroutine wrec_begin

  subroutine wrec_d(u, name, a)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    real(c_double), intent(in) :: a(:)
    character(len=32) :: nm
    nm = name
    write(u) nm, 1_c_int32_t, 0_c_int32_t, int(size(a), c_int64_t), a
  end subroutine wrec_d

  subroutine wrec_i(u, name, a)
    integer, intent(in) :: u
    character(len=*), intent(in) :: name
    integer(c_int32_t), intent(in) :: a(:)
    character(len=32) :: nm
    nm = name
    write(u) nm, 2_c_int32_t, 0_c_int32_t, int(size(a), c_int64_t), a
  end subroutine wrec_i

  subroutine wrec_end(u)
    integer, intent(in) :: u
    character(len=32) :: nm
    nm = "END"
    write(u) nm, 0_c_int32_t, 0_c_int32_t, 0_c_int64_t
  end subroutine wrec_end

end module mocdump_mod


program mocgpu_driver
  use, intrinsic :: iso_c_binding
  use mocdump_mod
  use mocgpu_mod
  implicit none
  character(len=1024) :: inPath, outPath, modelOut
  type(moc_tables) :: tb
  type(moc_grid_desc) :: gd
  type(moc_atomic) :: at
  type(moc_balance_tables) :: bt
  type(moc_emission_tables) :: et
  type(c_ptr) :: ctx
  real(c_double), allocatable, target :: Te(:), Ne(:), Hden(:), ionDen(:, :, :), J(:, :), &
       nPhot(:)
  real(c_double) :: totPercent
  integer :: nCells, nb, nElem, nst, nIter, k, u, nArg
  character(len=1) :: c
  ! run controls (iterateMC's loop: maxIterateMC, minConvergence, autoPackets)
  integer :: runMaxIter
  real(c_double) :: runMinConv, runConvInc, runPhotInc, runMaxPhot
  integer(c_int64_t) :: runSeed
  logical :: runAuto

  nArg = command_argument_count()
  call get_command_argument(1, inPath)
  if (trim(inPath) == "--namelist") then
     call get_command_argument(2, inPath)
     call get_command_argument(3, outPath)
     modelOut = ""
     if (nArg == 4 .and. trim(outPath) == "--write-model") then
        call get_command_argument(4, modelOut)
        outPath = ""
     else if (nArg /= 3) then
        print*, "usage: mocgpu_driver --namelist <namelist file> <output file>"
        print*, "       mocgpu_driver --namelist <namelist file> --write-model <model file>"
        stop
     end if
     call native_model(trim(inPath), trim(modelOut))
     if (len_trim(modelOut) > 0) stop
  else
     if (nArg /= 2) then
        print*, "usage: mocgpu_driver <model file> <output file>"
        stop
     end if
     call get_command_argument(2, outPath)
     call file_model(trim(inPath))
  end if

  call MCIterationDriver(ctx, nPhot, runMaxIter, runMinConv, runSeed, runAuto, runConvInc, &
       runPhotInc, runMaxPhot, nIter, totPercent)

  call get_field(ctx, MOC_F_TE, Te)
  call get_field(ctx, MOC_F_NE, Ne)
  call get_field(ctx, MOC_F_IONDEN, ionDen)
  call get_field(ctx, MOC_F_JSTE, J)
  call moc_destroy(ctx)

  open(newunit=u, file=trim(outPath), access="stream", form="unformatted", status="replace", &
       action="write")
  call wrec_begin(u)
  call wrec_d(u, "Te", Te)
  call wrec_d(u, "Ne", Ne)
  call wrec_d(u, "ionDen", reshape(ionDen, [size(ionDen)]))
  call wrec_d(u, "Jste", reshape(J, [size(J)]))
  call wrec_i(u, "nIterateMC", [int(nIter, c_int32_t)])
  call wrec_d(u, "totPercent", [totPercent])
  call wrec_d(u, "nPhotons", nPhot)
  call wrec_end(u)
  close(u)
  print*, "! mocgpu_driver: ", nIter, " iterations, ", totPercent, "% converged cells"

contains

  ! the model from a file of named records (fortran_io.write_model)
  subroutine file_model(path)
    character(len=*), intent(in) :: path
    call dump_read(path)
    print*, "! mocgpu_driver: ", c_string(moc_version())

    ! mesh and sources
    tb%nbins = ival("tables.nbins")
    tb%precision = ival("tables.precision")
    tb%nuArray = dptr("tables.nuArray")
    tb%widFlx = dptr("tables.widFlx")
    tb%ionEdge1 = dval("tables.ionEdge1")
    tb%lgCompton = ival("tables.lgCompton")
    tb%nAngles = ival("tables.nAngles")
    tb%KNsigmaT = dptr("tables.KNsigmaT")
    tb%KNsigmaCDF = dptr("tables.KNsigmaCDF")
    tb%Pcomp = dptr("tables.Pcomp")
    tb%nStars = ival("tables.nStars")
    tb%sourceCDF = dptr("tables.sourceCDF")
    tb%Lstar = dptr("tables.Lstar")
    tb%nPhotons = dptr("tables.nPhotons")
    tb%starPosition = dptr("tables.starPosition")
    tb%starIndex = iptr("tables.starIndex")
    call moc_check(ctx, moc_create(ctx, tb), "moc_create")

    ! the grid
    gd%nx = ival("grid.nx")
    gd%ny = ival("grid.ny")
    gd%nz = ival("grid.nz")
    gd%xAxis = dptr("grid.xAxis")
    gd%yAxis = dptr("grid.yAxis")
    gd%zAxis = dptr("grid.zAxis")
    gd%active = iptr("grid.active")
    gd%nCells = ival("grid.nCells")
    gd%lgSymmetricXYZ = ival("grid.lgSymmetricXYZ")
    gd%R_out = dval("grid.R_out")
    gd%lgPlaneIonization = ival("grid.lgPlaneIonization")
    call moc_check(ctx, moc_grid_attach(ctx, 1, gd), "moc_grid_attach")

    ! elements and cross-sections
    at%nElemUsed = ival("atomic.nElemUsed")
    at%elemZ = iptr("atomic.elemZ")
    at%abund = dptr("atomic.abund")
    at%nstages = ival("atomic.nstages")
    at%nSpecies = ival("atomic.nSpecies")
    at%speciesElem = iptr("atomic.speciesElem")
    at%speciesIon = iptr("atomic.speciesIon")
    at%sigma = dptr("atomic.sigma")
    at%brems = dptr("atomic.brems")
    call moc_check(ctx, moc_set_atomic(ctx, at), "moc_set_atomic")

    ! ionisation / thermal balance tables
    bt%maxLev = ival("balance.maxLev")
    bt%XHILimit = dval("balance.XHILimit")
    bt%nShells = ival("balance.nShells")
    bt%shellNum = iptr("balance.shellNum")
    bt%speciesShell = iptr("balance.speciesShell")
    bt%W = dptr("balance.W")
    bt%outShell = iptr("balance.outShell")
    bt%statW0 = iptr("balance.statW0")
    bt%statW1 = iptr("balance.statW1")
    bt%rrb = dptr("balance.rrb")
    bt%drb_c = dptr("balance.drb_c")
    bt%drb_e = dptr("balance.drb_e")
    bt%rrec = dptr("balance.rrec")
    bt%rnew = dptr("balance.rnew")
    bt%nDiel = ival("balance.nDiel")
    bt%diel = dptr("balance.diel")
    bt%nAldro = ival("balance.nAldro")
    bt%aldro = dptr("balance.aldro")
    bt%cionCF = dptr("balance.cionCF")
    bt%cionTherm = dptr("balance.cionTherm")
    bt%nAuger = iptr("balance.nAuger")
    bt%auger = dptr("balance.auger")
    bt%nChex = ival("balance.nChex")
    bt%chexZ = iptr("balance.chexZ")
    bt%chexIon = iptr("balance.chexIon")
    bt%chexFit = dptr("balance.chexFit")
    bt%nAtoms = ival("balance.nAtoms")
    bt%atomIdx = iptr("balance.atomIdx")
    bt%atomMeta = lptr("balance.atomMeta")
    bt%atomAtp = dptr("balance.atomAtp")
    bt%nLogT = lval("balance.nLogT")
    bt%nLev = lval("balance.nLev")
    bt%nA = lval("balance.nA")
    bt%nQ = lval("balance.nQ")
    bt%atomLogT = dptr("balance.atomLogT")
    bt%atomG = dptr("balance.atomG")
    bt%atomE = dptr("balance.atomE")
    bt%atomA = dptr("balance.atomA")
    bt%atomQom = dptr("balance.atomQom")
    bt%atomQ2 = dptr("balance.atomQ2")
    bt%atomRecKind = iptr("balance.atomRecKind")
    bt%atomRec = dptr("balance.atomRec")
    bt%atomRecValid = iptr("balance.atomRecValid")
    call moc_check(ctx, moc_set_balance(ctx, bt), "moc_set_balance")

    ! diffuse emission tables
    et%ntk = ival("emission.ntk")
    et%tkGamma = dptr("emission.tkGamma")
    do k = 1, 3
       write(c, '(I1)') k
       et%nlim(k) = ival("emission.nlim", k)
       et%edgeP(k) = iptr("emission.edgeP." // c)
       et%logGamma(k) = dptr("emission.logGamma." // c)
    end do
    et%HIPnuP = ival("emission.HIPnuP")
    et%HeIPnuP = ival("emission.HeIPnuP")
    et%HeIIPnuP = ival("emission.HeIIPnuP")
    et%xsHI = dptr("emission.xsHI")
    et%xsHeI = dptr("emission.xsHeI")
    et%xsHeII = dptr("emission.xsHeII")
    et%nFb = ival("emission.nFb")
    et%fbZ = iptr("emission.fbZ")
    et%fbIon = iptr("emission.fbIon")
    et%fbLo = iptr("emission.fbLo")
    et%fbHi = iptr("emission.fbHi")
    et%fbOff = lptr("emission.fbOff")
    et%nFbXs = lval("emission.nFbXs")
    et%fbXs = dptr("emission.fbXs")
    et%fbGratio = dptr("emission.fbGratio")
    et%he2Row = iptr("emission.he2Row")
    et%hei2phot = dptr("emission.hei2phot")
    et%nHlFiles = ival("emission.nHlFiles")
    et%hlMap = iptr("emission.hlMap")
    et%hlDens = dptr("emission.hlDens")
    et%hlLines = dptr("emission.hlLines")
    et%hbA = dptr("emission.hbA")
    et%hbB = dptr("emission.hbB")
    et%rbEdge = dptr("emission.rbEdge")
    et%r2aEdge = dptr("emission.r2aEdge")
    et%heiRec = dptr("emission.heiRec")
    et%heiiLyman = dptr("emission.heiiLyman")
    et%j2TsP = ival("emission.j2TsP")
    do k = 1, 4
       et%lymP(k) = ival("emission.lymP", k)
    end do
    call moc_check(ctx, moc_set_emission(ctx, et), "moc_set_emission")

    ! starting state in the reference's layout
    nCells = gd%nCells
    nb = tb%nbins
    nElem = at%nElemUsed
    nst = at%nstages
    allocate(Te(0:nCells), Ne(0:nCells), Hden(0:nCells), ionDen(0:nCells, nElem, nst), &
         J(0:nCells, nb), nPhot(tb%nStars))
    call copy_d("state.Te", Te)
    call copy_d("state.Ne", Ne)
    call copy_d("state.Hden", Hden)
    call copy_d3("state.ionDen", ionDen)
    call copy_d("tables.nPhotons", nPhot)
    call put_field(ctx, MOC_F_TE, Te)
    call put_field(ctx, MOC_F_NE, Ne)
    call put_field(ctx, MOC_F_HDEN, Hden)
    call put_field(ctx, MOC_F_IONDEN, ionDen)
    runMaxIter = int(ival("run.maxIterateMC"))
    runMinConv = dval("run.minConvergence")
    runSeed = lval("run.seed")
    runAuto = ival("run.lgAutoPackets") /= 0
    runConvInc = dval("run.convIncPercent")
    runPhotInc = dval("run.nPhotIncrease")
    runMaxPhot = dval("run.maxPhotons")
  end subroutine file_model

  ! the model built by the library from the reference's data/ files and a
  ! namelist of input-deck keywords (moc_model_build, moc_model_load); with
  ! writePath set, the model is written there and nothing runs on the device
  subroutine native_model(nmlPath, writePath)
    character(len=*), intent(in) :: nmlPath, writePath
    character(len=1024) :: dataDir
    character(len=16) :: precision
    integer :: elements(30), nstages, maxLev, nbins, nx, maxIterateMC, i, ios, nml
    real(c_double) :: logAbundances(30), XHILimit, nuMin, nuMax, TStellar, LStar, nPhotons, &
         plFraction, plLoKeV, plHiKeV, Rout, Hdensity, TeStart, H0Start, minConvergence, &
         convIncPercent, nPhotIncrease, maxPhotons
    integer(c_int64_t) :: seed
    logical :: strictMesh, lgCompton, lgAutoPackets
    character(kind=c_char, len=1025), target :: cDir
    integer(c_int32_t), allocatable, target :: zArr(:)
    real(c_double), allocatable, target :: abArr(:)
    type(moc_model_spec) :: spec
    type(c_ptr) :: mdl
    integer(c_int) :: rc
    namelist /model/ dataDir, elements, logAbundances, nstages, maxLev, XHILimit, nbins, nuMin, &
         nuMax, strictMesh, lgCompton, TStellar, LStar, nPhotons, plFraction, plLoKeV, plHiKeV, &
         nx, Rout, Hdensity, TeStart, H0Start, precision, maxIterateMC, minConvergence, seed, &
         lgAutoPackets, convIncPercent, nPhotIncrease, maxPhotons

    dataDir = "data"
    elements = 0
    logAbundances = 0.0_c_double
    nstages = 7
    maxLev = 27
    XHILimit = 0.05_c_double
    nbins = 1000
    nuMin = 1.0_c_double
    nuMax = 100.0_c_double
    strictMesh = .true.
    lgCompton = .true.
    TStellar = 1.e5_c_double
    LStar = 10.0_c_double
    nPhotons = 1.e6_c_double
    plFraction = 0.0_c_double
    plLoKeV = 0.01_c_double
    plHiKeV = 10.0_c_double
    nx = 16
    Rout = 1.e18_c_double
    Hdensity = 100.0_c_double
    TeStart = 1.e4_c_double
    H0Start = 1.e-5_c_double
    precision = "fp64"
    maxIterateMC = 5
    minConvergence = 100.0_c_double
    seed = 20261203_c_int64_t
    lgAutoPackets = .false.
    convIncPercent = 0.0_c_double
    nPhotIncrease = 2.0_c_double
    maxPhotons = 0.0_c_double
    open(newunit=nml, file=nmlPath, status="old", action="read", iostat=ios)
    if (ios /= 0) then
       print*, "! mocgpu_driver: can't open namelist file: ", nmlPath
       stop
    end if
    read(nml, nml=model, iostat=ios)
    close(nml)
    if (ios /= 0) then
       print*, "! mocgpu_driver: can't read the &model group of ", nmlPath
       stop
    end if
    nElem = count(elements > 0)
    if (nElem == 0 .or. any(elements(nElem + 1:) /= 0)) then
       print*, "! mocgpu_driver: elements must list the atomic numbers first"
       stop
    end if
    allocate(zArr(nElem), abArr(nElem))
    zArr = int(elements(1:nElem), c_int32_t)
    do i = 1, nElem
       abArr(i) = 10.0_c_double ** logAbundances(i)
    end do
    cDir = trim(dataDir) // c_null_char
    spec%dataDir = c_loc(cDir)
    select case (trim(precision))
    case ("fp64")
       spec%precision = MOC_PREC_FP64
    case ("mixed")
       spec%precision = MOC_PREC_MIXED
    case default
       print*, "! mocgpu_driver: precision must be 'fp64' or 'mixed'"
       stop
    end select
    spec%nElem = nElem
    spec%elemZ = c_loc(zArr)
    spec%abund = c_loc(abArr)
    spec%nstages = nstages
    spec%maxLev = maxLev
    spec%XHILimit = XHILimit
    spec%nbins = nbins
    spec%nuMin = nuMin
    spec%nuMax = nuMax
    spec%strictMesh = merge(1, 0, strictMesh)
    spec%lgCompton = merge(1, 0, lgCompton)
    spec%Tstellar = TStellar
    spec%LStar = LStar
    spec%nPhotons = nPhotons
    spec%plFraction = plFraction
    spec%plLoKeV = plLoKeV
    spec%plHiKeV = plHiKeV
    spec%nx = nx
    spec%Rout = Rout
    spec%Hdens = Hdensity
    spec%hden = c_null_ptr
    spec%TeStart = TeStart
    spec%H0in = H0Start
    rc = moc_model_build(spec, mdl)
    if (rc /= MOC_OK) then
       print*, "! mocgpu_driver: ", c_string(moc_model_last_error())
       stop 1
    end if
    if (len(writePath) > 0) then
       rc = moc_model_write(mdl, trim(writePath) // c_null_char)
       if (rc /= MOC_OK) then
          print*, "! mocgpu_driver: ", c_string(moc_model_last_error())
          stop 1
       end if
       call moc_model_free(mdl)
       print*, "! mocgpu_driver: model written to ", writePath
       return
    end if
    print*, "! mocgpu_driver: ", c_string(moc_version())
    rc = moc_model_load(ctx, mdl)
    if (rc /= MOC_OK) then
       print*, "! mocgpu_driver: ", c_string(moc_model_last_error())
       stop 1
    end if
    rc = moc_model_get(mdl, tb, gd, at, bt, et)
    nCells = gd%nCells
    nb = tb%nbins
    nElem = at%nElemUsed
    nst = at%nstages
    call moc_model_free(mdl)
    allocate(Te(0:nCells), Ne(0:nCells), Hden(0:nCells), ionDen(0:nCells, nElem, nst), &
         J(0:nCells, nb), nPhot(1))
    nPhot(1) = nPhotons
    runMaxIter = maxIterateMC
    runMinConv = minConvergence
    runSeed = seed
    runAuto = lgAutoPackets
    runConvInc = convIncPercent
    runPhotInc = nPhotIncrease
    runMaxPhot = maxPhotons
  end subroutine native_model

  subroutine copy_d(name, a)
    character(len=*), intent(in) :: name
    real(c_double), intent(out), target :: a(:)
    real(c_double), pointer :: p(:)
    type(c_ptr) :: q
    q = dptr(name)
    if (.not. c_associated(q)) then
       print*, "! mocgpu_driver: record missing: ", name
       stop
    end if
    call c_f_pointer(q, p, [size(a)])
    a = p
  end subroutine copy_d

  subroutine copy_d3(name, a)
    character(len=*), intent(in) :: name
    real(c_double), intent(out), target :: a(:, :, :)
    real(c_double), pointer :: p(:, :, :)
    type(c_ptr) :: q
    q = dptr(name)
    if (.not. c_associated(q)) then
       print*, "! mocgpu_driver: record missing: ", name
       stop
    end if
    call c_f_pointer(q, p, shape(a))
    a = p
  end subroutine copy_d3

end program mocgpu_driver
