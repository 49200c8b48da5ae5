! mocgpu_mod -- the Fortran side of libmocgpu: ISO_C_BINDING types and
! interfaces for every entry point of include/mocgpu.h, and the MOCASSIN
! drivers rebuilt on them with the reference's names and control flow:
!
!   ionizationDriver   <- ionization_mod.f90:26   (every cell at once)
!   emissionDriver     <- emission_mod.f90:36     (every cell at once)
!   energyPacketDriver <- photon_mod.f90:28       (pathSegment runs on the device)
!   updateCell         <- update_mod.f90:15       (every cell at once)
!   MCIterationDriver  <- iteration_mod.f90:17    (the iterateMC loop: convergence
!                                                  percentage, minConvergence,
!                                                  maxIterateMC, autoPackets)
!
! The model set-up of include/mocmodel.h (moc_model_build and friends) reads
! the reference's data/ files in the library: a Fortran host then needs no
! other program to build the tables (mocgpu_driver --namelist).
!
! The grid arrays stay in the reference's layout on the Fortran side:
! (0:nCells) per cell, (0:nCells, 1:nbins) per cell and frequency (cell index
! fastest) and (0:nCells, nElementsUsed, nstages) for ionDen; put_field /
! get_field move them through moc_upload / moc_download, which transpose on
! the host.  Every nonzero return code stops the run with the library's
! message, as the reference's print* + stop does.
module mocgpu_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MOC_OK = 0
  integer(c_int), parameter, public :: MOC_PREC_FP64 = 0, MOC_PREC_MIXED = 1
  integer(c_int), parameter, public :: MOC_DT_F32 = 0, MOC_DT_F64 = 1, MOC_DT_I32 = 2
  integer(c_int), parameter, public :: MOC_LAYOUT_FORTRAN = 0, MOC_LAYOUT_CELLROW = 1
  integer(c_int), parameter, public :: MOC_F_OPACITY = 1, MOC_F_JSTE = 2, MOC_F_RECPDF = 3, &
       MOC_F_TOTALLINES = 4, MOC_F_NE = 5, MOC_F_TE = 6, MOC_F_HDEN = 7, MOC_F_ESCAPED = 8, &
       MOC_F_IONDEN = 9, MOC_F_CONVERGED = 10, MOC_F_BLACK = 11, MOC_F_EMISSION = 12, &
       MOC_F_NDUST = 13, MOC_F_TDUST = 14, MOC_F_DUSTPDF = 15, MOC_F_ABSOPAC = 16, &
       MOC_F_SCAOPAC = 17, MOC_F_GRAINCOLL = 18, MOC_F_GRAINPE = 19

  type, bind(C), public :: moc_tables
     integer(c_int32_t) :: nbins, precision
     type(c_ptr) :: nuArray, widFlx
     real(c_double) :: ionEdge1
     integer(c_int32_t) :: lgCompton, nAngles
     type(c_ptr) :: KNsigmaT, KNsigmaCDF, Pcomp
     integer(c_int32_t) :: nStars
     type(c_ptr) :: sourceCDF, Lstar, nPhotons, starPosition, starIndex
  end type moc_tables

  type, bind(C), public :: moc_grid_desc
     integer(c_int32_t) :: nx, ny, nz
     type(c_ptr) :: xAxis, yAxis, zAxis, active
     integer(c_int32_t) :: nCells, lgSymmetricXYZ
     real(c_double) :: R_out
     integer(c_int32_t) :: lgPlaneIonization
  end type moc_grid_desc

  type, bind(C), public :: moc_atomic
     integer(c_int32_t) :: nElemUsed
     type(c_ptr) :: elemZ, abund
     integer(c_int32_t) :: nstages, nSpecies
     type(c_ptr) :: speciesElem, speciesIon, sigma, brems
  end type moc_atomic

  type, bind(C), public :: moc_balance_tables
     integer(c_int32_t) :: maxLev
     real(c_double) :: XHILimit
     integer(c_int32_t) :: nShells
     type(c_ptr) :: shellNum, speciesShell, W, outShell, statW0, statW1, rrb, drb_c, drb_e, &
          rrec, rnew
     integer(c_int32_t) :: nDiel
     type(c_ptr) :: diel
     integer(c_int32_t) :: nAldro
     type(c_ptr) :: aldro, cionCF, cionTherm, nAuger, auger
     integer(c_int32_t) :: nChex
     type(c_ptr) :: chexZ, chexIon, chexFit
     integer(c_int32_t) :: nAtoms
     type(c_ptr) :: atomIdx, atomMeta, atomAtp
     integer(c_int64_t) :: nLogT, nLev, nA, nQ
     type(c_ptr) :: atomLogT, atomG, atomE, atomA, atomQom, atomQ2, atomRecKind, atomRec, &
          atomRecValid
  end type moc_balance_tables

  type, bind(C), public :: moc_emission_tables
     integer(c_int32_t) :: ntk
     type(c_ptr) :: tkGamma
     integer(c_int32_t) :: nlim(3)
     type(c_ptr) :: edgeP(3), logGamma(3)
     integer(c_int32_t) :: HIPnuP, HeIPnuP, HeIIPnuP
     type(c_ptr) :: xsHI, xsHeI, xsHeII
     integer(c_int32_t) :: nFb
     type(c_ptr) :: fbZ, fbIon, fbLo, fbHi, fbOff
     integer(c_int64_t) :: nFbXs
     type(c_ptr) :: fbXs, fbGratio, he2Row, hei2phot
     integer(c_int32_t) :: nHlFiles
     type(c_ptr) :: hlMap, hlDens, hlLines, hbA, hbB, rbEdge, r2aEdge, heiRec, heiiLyman
     integer(c_int32_t) :: j2TsP, lymP(4)
  end type moc_emission_tables

  ! grain tables (dustDriver); Tdust is (0:nSpecies, 0:nSizes, 0:nCells)
  type, bind(C), public :: moc_dust_tables
     integer(c_int32_t) :: nSpecies, nSizes, nTemps, lgGas, lgPhotoelectric
     type(c_ptr) :: grainAbun, grainWeight, TdustSublime, Cabs, Csca, emIntegral
     real(c_double) :: XHILimit
     type(c_ptr) :: grainRadius, grainVn, MsurfAtom, ipRyd
     ! quantum heating (qHeat): lgQHeat, nTbins, minaQHeat, minConvQHeat, Tmax, grainKind
     integer(c_int32_t) :: lgQHeat = 0, nTbins = 0
     real(c_double) :: minaQHeat = 0d0, minConvQHeat = 0d0, Tmax = 0d0
     type(c_ptr) :: grainKind = c_null_ptr
     ! lgEquivalentTau: iteration 1 without the grains' opacity
     integer(c_int32_t) :: lgEquivalentTau = 0
  end type moc_dust_tables

  type, bind(C), public :: moc_transport_stats
     real(c_double) :: Qphot, segments, emissions, absorbed, scattered, escaped, lines, lost, &
          trapped, kernel_ms
  end type moc_transport_stats

  type, bind(C), public :: moc_update_stats
     real(c_double) :: noHit, noIonBal, noTeBal, converged, kernel_ms
  end type moc_update_stats

  type, bind(C), public :: moc_iteration_stats
     real(c_double) :: convPercent, noHitPercent, noIonBalPercent, noTeBalPercent, packets, &
          opacity_ms, emission_ms, transport_ms, update_ms, segments, emissions, escaped, &
          iteration_ms
  end type moc_iteration_stats

  ! launch parameters of the transport (moc_set_tuning); graph is set by moc_get_tuning
  type, bind(C), public :: moc_tuning
     integer(c_int32_t) :: emitBatch, segsPerTrip, transportWaves, graph, hotJ
  end type moc_tuning

  ! one gas model for moc_model_build (include/mocmodel.h)
  type, bind(C), public :: moc_model_spec
     type(c_ptr) :: dataDir
     integer(c_int32_t) :: precision, nElem
     type(c_ptr) :: elemZ, abund
     integer(c_int32_t) :: nstages, maxLev
     real(c_double) :: XHILimit
     integer(c_int32_t) :: nbins
     real(c_double) :: nuMin, nuMax
     integer(c_int32_t) :: strictMesh, lgCompton
     real(c_double) :: Tstellar, LStar, nPhotons, plFraction, plLoKeV, plHiKeV
     integer(c_int32_t) :: nx
     real(c_double) :: Rout, Hdens
     type(c_ptr) :: hden
     real(c_double) :: TeStart, H0in
  end type moc_model_spec

  interface
     integer(c_int) function moc_model_build(spec, model) bind(C, name="moc_model_build")
       import :: c_int, c_ptr, moc_model_spec
       type(moc_model_spec), intent(in) :: spec
       type(c_ptr), intent(out) :: model
     end function moc_model_build
     subroutine moc_model_free(model) bind(C, name="moc_model_free")
       import :: c_ptr
       type(c_ptr), value :: model
     end subroutine moc_model_free
     type(c_ptr) function moc_model_last_error() bind(C, name="moc_model_last_error")
       import :: c_ptr
     end function moc_model_last_error
     integer(c_int) function moc_model_get(model, tables, grid, atomic, balance, emission) &
          bind(C, name="moc_model_get")
       import :: c_int, c_ptr, moc_tables, moc_grid_desc, moc_atomic, moc_balance_tables, &
            moc_emission_tables
       type(c_ptr), value :: model
       type(moc_tables), intent(out) :: tables
       type(moc_grid_desc), intent(out) :: grid
       type(moc_atomic), intent(out) :: atomic
       type(moc_balance_tables), intent(out) :: balance
       type(moc_emission_tables), intent(out) :: emission
     end function moc_model_get
     integer(c_int) function moc_model_state(model, nCells, Te, Ne, Hden, ionDen) &
          bind(C, name="moc_model_state")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: model
       integer(c_int64_t), intent(out) :: nCells
       type(c_ptr), intent(out) :: Te, Ne, Hden, ionDen
     end function moc_model_state
     integer(c_int) function moc_model_load(ctx, model) bind(C, name="moc_model_load")
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: ctx
       type(c_ptr), value :: model
     end function moc_model_load
     integer(c_int) function moc_model_write(model, path) bind(C, name="moc_model_write")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: model
       character(kind=c_char), intent(in) :: path(*)
     end function moc_model_write
  end interface

  interface
     integer(c_int) function moc_create(ctx, tables) bind(C, name="moc_create")
       import :: c_int, c_ptr, moc_tables
       type(c_ptr), intent(out) :: ctx
       type(moc_tables), intent(in) :: tables
     end function moc_create
     subroutine moc_destroy(ctx) bind(C, name="moc_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine moc_destroy
     type(c_ptr) function moc_last_error(ctx) bind(C, name="moc_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function moc_last_error
     type(c_ptr) function moc_version() bind(C, name="moc_version")
       import :: c_ptr
     end function moc_version
     integer(c_int) function moc_grid_attach(ctx, ig, grid) bind(C, name="moc_grid_attach")
       import :: c_int, c_ptr, moc_grid_desc
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
       type(moc_grid_desc), intent(in) :: grid
     end function moc_grid_attach
     ! a batch of independent grids (star k lights grid k; cells numbered grid
     ! after grid)
     integer(c_int) function moc_grid_attach_batch(ctx, nGrids, grids) &
          bind(C, name="moc_grid_attach_batch")
       import :: c_int, c_ptr, moc_grid_desc
       type(c_ptr), value :: ctx
       integer(c_int), value :: nGrids
       type(moc_grid_desc), intent(in) :: grids(*)
     end function moc_grid_attach_batch
     ! a mother grid and its nested sub-grids (multiGrids)
     integer(c_int) function moc_grid_attach_nested(ctx, nGrids, grids) &
          bind(C, name="moc_grid_attach_nested")
       import :: c_int, c_ptr, moc_grid_desc
       type(c_ptr), value :: ctx
       integer(c_int), value :: nGrids
       type(moc_grid_desc), intent(in) :: grids(*)
     end function moc_grid_attach_nested
     integer(c_int) function moc_upload(ctx, field, ig, host, count, dtype, layout) &
          bind(C, name="moc_upload")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, host
       integer(c_int), value :: field, ig, dtype, layout
       integer(c_int64_t), value :: count
     end function moc_upload
     integer(c_int) function moc_download(ctx, field, ig, host, count, dtype, layout) &
          bind(C, name="moc_download")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx, host
       integer(c_int), value :: field, ig, dtype, layout
       integer(c_int64_t), value :: count
     end function moc_download
     integer(c_int) function moc_set_atomic(ctx, atomic) bind(C, name="moc_set_atomic")
       import :: c_int, c_ptr, moc_atomic
       type(c_ptr), value :: ctx
       type(moc_atomic), intent(in) :: atomic
     end function moc_set_atomic
     integer(c_int) function moc_opacity(ctx, ig) bind(C, name="moc_opacity")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
     end function moc_opacity
     integer(c_int) function moc_set_balance(ctx, tables) bind(C, name="moc_set_balance")
       import :: c_int, c_ptr, moc_balance_tables
       type(c_ptr), value :: ctx
       type(moc_balance_tables), intent(in) :: tables
     end function moc_set_balance
     integer(c_int) function moc_set_emission(ctx, tables) bind(C, name="moc_set_emission")
       import :: c_int, c_ptr, moc_emission_tables
       type(c_ptr), value :: ctx
       type(moc_emission_tables), intent(in) :: tables
     end function moc_set_emission
     integer(c_int) function moc_emission(ctx, ig, kernel_ms) bind(C, name="moc_emission")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
       real(c_float), intent(out) :: kernel_ms
     end function moc_emission
     integer(c_int) function moc_set_packets(ctx, iStar, n) bind(C, name="moc_set_packets")
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int), value :: iStar
       integer(c_int64_t), value :: n
     end function moc_set_packets
     integer(c_int) function moc_iterate(ctx, ig, nIterateMC, seed, stats) &
          bind(C, name="moc_iterate")
       import :: c_int, c_ptr, c_int64_t, moc_iteration_stats
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig, nIterateMC
       integer(c_int64_t), value :: seed
       type(moc_iteration_stats), intent(out) :: stats
     end function moc_iterate
     integer(c_int) function moc_update(ctx, ig, stats) bind(C, name="moc_update")
       import :: c_int, c_ptr, moc_update_stats
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
       type(moc_update_stats), intent(out) :: stats
     end function moc_update
     integer(c_int) function moc_reset_tallies(ctx, ig) bind(C, name="moc_reset_tallies")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
     end function moc_reset_tallies
     integer(c_int) function moc_transport(ctx, iStar, n, seed, ctr_base, stats) &
          bind(C, name="moc_transport")
       import :: c_int, c_ptr, c_int64_t, c_int32_t, moc_transport_stats
       type(c_ptr), value :: ctx
       integer(c_int), value :: iStar
       integer(c_int64_t), value :: n, seed
       integer(c_int32_t), value :: ctr_base
       type(moc_transport_stats), intent(out) :: stats
     end function moc_transport
     integer(c_int) function moc_normalize_j(ctx, ig) bind(C, name="moc_normalize_j")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
     end function moc_normalize_j
     integer(c_int) function moc_set_dust(ctx, tables) bind(C, name="moc_set_dust")
       import :: c_int, c_ptr, moc_dust_tables
       type(c_ptr), value :: ctx
       type(moc_dust_tables), intent(in) :: tables
     end function moc_set_dust
     integer(c_int) function moc_dust_opacity(ctx, ig) bind(C, name="moc_dust_opacity")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
     end function moc_dust_opacity
     integer(c_int) function moc_dust_emission(ctx, ig) bind(C, name="moc_dust_emission")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
     end function moc_dust_emission
     integer(c_int) function moc_dust_temperature(ctx, ig, stats) &
          bind(C, name="moc_dust_temperature")
       import :: c_int, c_ptr, moc_update_stats
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
       type(moc_update_stats), intent(out) :: stats
     end function moc_dust_temperature
     ! outputGas volume averages: TeVol, ionDenVol (nElemUsed, nstages)
     integer(c_int) function moc_volume_averages(ctx, ig, TeVol, ionDenVol) &
          bind(C, name="moc_volume_averages")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
       real(c_double), intent(out) :: TeVol(*), ionDenVol(*)
     end function moc_volume_averages
     ! outputGas line luminosities (lines(j, i, k): ion k, upper level j,
     ! lower level i) and Hbeta; lines as c_null_ptr returns the sizes only
     integer(c_int) function moc_line_luminosities(ctx, ig, nIons, nLev, ionZ, ionStage, lines, &
          count, Hbeta) bind(C, name="moc_line_luminosities")
       import :: c_int, c_ptr, c_int64_t, c_double
       type(c_ptr), value :: ctx, ionZ, ionStage, lines
       integer(c_int), value :: ig
       integer(c_int), intent(out) :: nIons, nLev
       integer(c_int64_t), value :: count
       real(c_double), intent(out) :: Hbeta
     end function moc_line_luminosities
     ! writeSED, all-angle column
     integer(c_int) function moc_sed(ctx, ig, sed, totalE) bind(C, name="moc_sed")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ig
       real(c_double), intent(out) :: sed(*), totalE
     end function moc_sed
     ! replay of moc_iterate as one captured graph (no counterpart in the reference)
     integer(c_int) function moc_set_graph(ctx, enable) bind(C, name="moc_set_graph")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: enable
     end function moc_set_graph
     ! launch parameters of the transport (no counterpart in the reference)
     integer(c_int) function moc_set_tuning(ctx, t) bind(C, name="moc_set_tuning")
       import :: c_int, c_ptr, moc_tuning
       type(c_ptr), value :: ctx
       type(moc_tuning), intent(in) :: t
     end function moc_set_tuning
     integer(c_int) function moc_get_tuning(ctx, t) bind(C, name="moc_get_tuning")
       import :: c_int, c_ptr, moc_tuning
       type(c_ptr), value :: ctx
       type(moc_tuning), intent(out) :: t
     end function moc_get_tuning
     ! the iteration state that switches quantum heating on (nIterateMC, convPercent
     ! of the reference's module variables)
     integer(c_int) function moc_set_qheat_state(ctx, nIterateMC, convPercent) &
          bind(C, name="moc_set_qheat_state")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nIterateMC
       real(c_double), value :: convPercent
     end function moc_set_qheat_state
  end interface

  public :: moc_set_graph, moc_set_tuning, moc_get_tuning, moc_set_qheat_state
  public :: moc_model_build, moc_model_free, moc_model_last_error, moc_model_get, &
       moc_model_state, moc_model_load, moc_model_write
  public :: moc_create, moc_destroy, moc_last_error, moc_version, moc_grid_attach, moc_upload, &
       moc_download, moc_set_atomic, moc_opacity, moc_set_balance, moc_set_emission, &
       moc_emission, moc_set_packets, moc_iterate, moc_update, moc_reset_tallies, moc_transport, &
       moc_normalize_j, moc_set_dust, moc_dust_opacity, moc_dust_emission, moc_dust_temperature, &
       moc_grid_attach_batch, moc_grid_attach_nested, moc_volume_averages, moc_line_luminosities, moc_sed
  public :: moc_check, c_string, put_field, get_field, ionizationDriver, emissionDriver, &
       energyPacketDriver, updateCell, MCIterationDriver, dustDriver, getDustT

  ! real(8) arrays, or default-real ones as the reference's grid_type holds
  interface put_field
     module procedure put_field_1d, put_field_2d, put_field_3d, put_field_1s, put_field_2s, &
          put_field_3s
  end interface put_field
  interface get_field
     module procedure get_field_1d, get_field_2d, get_field_3d, get_field_1s, get_field_2s, &
          get_field_3s
  end interface get_field

contains

  ! a C string as a Fortran character variable
  function c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: a(:)
    integer :: n
    interface
       integer(c_size_t) function strlen(q) bind(C, name="strlen")
         import :: c_size_t, c_ptr
         type(c_ptr), value :: q
       end function strlen
    end interface
    if (.not. c_associated(p)) then
       s = ""
       return
    end if
    n = int(strlen(p))
    call c_f_pointer(p, a, [n])
    allocate(character(len=n) :: s)
    s = transfer(a, s)
  end function c_string

  ! the reference's error convention: report and stop
  subroutine moc_check(ctx, rc, where)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= MOC_OK) then
       print*, "! ", where, ": ", c_string(moc_last_error(ctx)), " (rc=", rc, ")"
       stop
    end if
  end subroutine moc_check

  subroutine put_field_1d(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_double), intent(in), target, contiguous :: a(:)
    call moc_check(ctx, moc_upload(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F64, MOC_LAYOUT_FORTRAN), "put_field")
  end subroutine put_field_1d

  subroutine put_field_2d(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_double), intent(in), target, contiguous :: a(:, :)
    call moc_check(ctx, moc_upload(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F64, MOC_LAYOUT_FORTRAN), "put_field")
  end subroutine put_field_2d

  subroutine put_field_3d(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_double), intent(in), target, contiguous :: a(:, :, :)
    call moc_check(ctx, moc_upload(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F64, MOC_LAYOUT_FORTRAN), "put_field")
  end subroutine put_field_3d

  subroutine get_field_1d(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_double), intent(inout), target, contiguous :: a(:)
    call moc_check(ctx, moc_download(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F64, MOC_LAYOUT_FORTRAN), "get_field")
  end subroutine get_field_1d

  subroutine get_field_2d(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_double), intent(inout), target, contiguous :: a(:, :)
    call moc_check(ctx, moc_download(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F64, MOC_LAYOUT_FORTRAN), "get_field")
  end subroutine get_field_2d

  subroutine get_field_3d(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_double), intent(inout), target, contiguous :: a(:, :, :)
    call moc_check(ctx, moc_download(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F64, MOC_LAYOUT_FORTRAN), "get_field")
  end subroutine get_field_3d

  subroutine put_field_1s(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_float), intent(in), target, contiguous :: a(:)
    call moc_check(ctx, moc_upload(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F32, MOC_LAYOUT_FORTRAN), "put_field")
  end subroutine put_field_1s

  subroutine put_field_2s(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_float), intent(in), target, contiguous :: a(:, :)
    call moc_check(ctx, moc_upload(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F32, MOC_LAYOUT_FORTRAN), "put_field")
  end subroutine put_field_2s

  subroutine put_field_3s(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_float), intent(in), target, contiguous :: a(:, :, :)
    call moc_check(ctx, moc_upload(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F32, MOC_LAYOUT_FORTRAN), "put_field")
  end subroutine put_field_3s

  subroutine get_field_1s(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_float), intent(inout), target, contiguous :: a(:)
    call moc_check(ctx, moc_download(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F32, MOC_LAYOUT_FORTRAN), "get_field")
  end subroutine get_field_1s

  subroutine get_field_2s(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_float), intent(inout), target, contiguous :: a(:, :)
    call moc_check(ctx, moc_download(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F32, MOC_LAYOUT_FORTRAN), "get_field")
  end subroutine get_field_2s

  subroutine get_field_3s(ctx, field, a)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: field
    real(c_float), intent(inout), target, contiguous :: a(:, :, :)
    call moc_check(ctx, moc_download(ctx, field, 1, c_loc(a), int(size(a), c_int64_t), &
         MOC_DT_F32, MOC_LAYOUT_FORTRAN), "get_field")
  end subroutine get_field_3s

  ! opacities of every cell (ionization_mod.f90:26, called per cell from
  ! iteration_mod.f90:148)
  subroutine ionizationDriver(ctx)
    type(c_ptr), intent(in) :: ctx
    call moc_check(ctx, moc_opacity(ctx, 1), "ionizationDriver")
  end subroutine ionizationDriver

  ! diffuse emission CDFs and line fractions of every cell (emission_mod.f90:36)
  subroutine emissionDriver(ctx)
    type(c_ptr), intent(in) :: ctx
    real(c_float) :: ms
    call moc_check(ctx, moc_emission(ctx, 1, ms), "emissionDriver")
  end subroutine emissionDriver

  ! n packets of star iStar through the grid (photon_mod.f90:28); the J and
  ! escape tallies accumulate on the device until moc_reset_tallies
  subroutine energyPacketDriver(ctx, iStar, n, seed, nIterateMC, stats)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: iStar, nIterateMC
    integer(c_int64_t), intent(in) :: n, seed
    type(moc_transport_stats), intent(out) :: stats
    call moc_check(ctx, moc_transport(ctx, iStar, n, seed, &
         int(nIterateMC * 256 + iStar, c_int32_t), stats), "energyPacketDriver")
  end subroutine energyPacketDriver

  ! the balance of every cell (update_mod.f90:15) from the normalised J
  subroutine updateCell(ctx, stats)
    type(c_ptr), intent(in) :: ctx
    type(moc_update_stats), intent(out) :: stats
    call moc_check(ctx, moc_update(ctx, 1, stats), "updateCell")
  end subroutine updateCell

  ! dustDriver (dust_mod.f90:14-186): the grain tables, built on the host
  ! (cross-sections, weights, Kirchhoff integrals), go to the device; then
  ! put_field(ctx, MOC_F_NDUST, grid(1)%Ndust).  The dust opacity, the dust
  ! emission CDF and getDustT run inside moc_iterate and updateCell.
  subroutine dustDriver(ctx, tables)
    type(c_ptr), intent(in) :: ctx
    type(moc_dust_tables), intent(in) :: tables
    call moc_check(ctx, moc_set_dust(ctx, tables), "dustDriver")
  end subroutine dustDriver

  ! getDustT (update_mod.f90:3346-3453) over every cell on its own
  subroutine getDustT(ctx, stats)
    type(c_ptr), intent(in) :: ctx
    type(moc_update_stats), intent(out) :: stats
    call moc_check(ctx, moc_dust_temperature(ctx, 1, stats), "getDustT")
  end subroutine getDustT

  ! The iterateMC loop (iteration_mod.f90:32-1497 without its output files):
  ! one device iteration per pass, the converged-cell percentage against
  ! minConvergence and maxIterateMC, and with lgAutoPackets the packet count
  ! of every star raised by nPhotIncrease when the percentage grew by no more
  ! than convIncPercent (:1412-1421).  nPhotons is updated in place.
  subroutine MCIterationDriver(ctx, nPhotons, maxIterateMC, minConvergence, seed, &
       lgAutoPackets, convIncPercent, nPhotIncrease, maxPhotons, nIterateMC, totPercent)
    type(c_ptr), intent(in) :: ctx
    real(c_double), intent(inout) :: nPhotons(:)
    integer, intent(in) :: maxIterateMC
    real(c_double), intent(in) :: minConvergence, convIncPercent, nPhotIncrease, maxPhotons
    integer(c_int64_t), intent(in) :: seed
    logical, intent(in) :: lgAutoPackets
    integer, intent(out) :: nIterateMC
    real(c_double), intent(out) :: totPercent
    type(moc_iteration_stats) :: st
    real(c_double) :: totPercentOld, nPhotonsTot
    integer :: iStar
    totPercentOld = 0.d0
    nIterateMC = 1
    do
       call moc_check(ctx, moc_iterate(ctx, 1, int(nIterateMC, c_int), seed, st), &
            "MCIterationDriver")
       totPercent = st%convPercent
       print '(A,I4,A,F8.3,A,F8.3,A,F8.3,A,F8.3,A)', " ! iterateMC: iteration ", nIterateMC, &
            ": ", totPercent, "% converged cells, ", st%noHitPercent, "% no hit, ", &
            st%noIonBalPercent, "% ion bal not reached, ", st%noTeBalPercent, &
            "% Te bal not reached"
       nPhotonsTot = nPhotons(1) + sum(nPhotons)
       if (lgAutoPackets .and. nIterateMC > 1 .and. totPercent < 95.d0 .and. &
            nPhotonsTot < maxPhotons .and. totPercentOld > 0.d0) then
          if ((totPercent - totPercentOld) / totPercentOld <= convIncPercent) then
             nPhotons = nPhotons * nPhotIncrease
             do iStar = 1, size(nPhotons)
                call moc_check(ctx, moc_set_packets(ctx, int(iStar, c_int), &
                     int(nPhotons(iStar), c_int64_t)), "MCIterationDriver")
             end do
             print*, "! iterateMC: nPhotons increased to ", nPhotons
          end if
       end if
       totPercentOld = totPercent
       if (totPercent >= minConvergence .or. nIterateMC >= maxIterateMC) exit
       nIterateMC = nIterateMC + 1
    end do
  end subroutine MCIterationDriver

end module mocgpu_mod
