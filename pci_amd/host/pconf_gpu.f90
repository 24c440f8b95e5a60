! pconf_gpu -- drop-in for `pconf` on one MI355X.
!
! Reads CONF.INP, ci.in, CONF.DAT, CONF.INT, CONF.GNT from the working
! directory like pconf (src/pconf.F90:87-143), builds H and J^2 on the GPU
! (FormH/RedistributeHamCSR/FormJ, src/pconf.F90:201-238), writes
! pCONF.HIJ/pCONF.JJJ when Kw=1 or reads pCONF.HIJ when Kl=1, runs the
! Davidson procedure (Diag4) on the GPU and writes CONF.DET, CONF.XIJ and
! the energy table to CONF.RES and stdout.
!
! With kCSF = 1 in ci.in the solve runs in the CSF basis (src/pconf.F90:
! 146-149, 211-214, 247-255): configuration classes before the determinant
! list, the CSFs and H in their basis on the GPU, Davidson there, CONF.WFS
! with the CSF-basis vectors and CONF.XIJ with the determinant-basis ones.
!
! Environment: PCI_DVD_FLAGS (bit 1: FP32 H with FP64 refinement, bit 2:
! learned correction) selects the fast variants; the default is the plain
! FP64 method.  With bit 2, PCI_MLP_WEIGHTS names the network file (raw
! little-endian FP32, pci_gpu_mlp_nparams values).  PCI_DEVICE selects the
! GPU ordinal.
Program pconf_gpu
    Use, Intrinsic :: iso_c_binding
    Use, Intrinsic :: iso_fortran_env, Only : dp => real64, int64, int32
    Use pci_gpu_api
    Use pconf_host
    Implicit None
    Type(c_ptr) :: ctx
    Type(pci_dvd_params) :: prm
    Type(pci_dvd_stats) :: st
    Integer(c_int) :: rc
    Integer(c_int64_t) :: numh, numzero, numj
    Real(c_double) :: hmin
    Integer(int64), Allocatable :: rowptr(:)
    Integer(int32), Allocatable :: col(:)
    Real(dp), Allocatable :: val(:), tk(:), tj(:), vecs(:, :)
    Integer(int64) :: t0, t1, rate
    Integer :: flags, dev, ios
    Character(Len=32) :: env
    Character(Len=512) :: wfile
    Integer(c_int64_t) :: nw
    Real(c_float), Allocatable :: w(:)
    Integer :: jt(256), njd
    Logical :: have_hij, legacy
    Integer(c_int32_t) :: ncsf
    Integer(c_int64_t) :: nccj
    Integer(int32), Allocatable :: ndcs(:)
    Real(dp), Allocatable :: dvecs(:, :)
    Integer :: nbas, nread, ndpt
    Real(dp), Allocatable, Target :: tk0(:), x0(:, :)
    Type(c_ptr) :: ptk0, px0

    Call System_Clock(t0, rate)
    Open(unit=11, file='CONF.RES', status='UNKNOWN')
    Write(*, '(4X,"Program pconf_gpu (MI355X) with double precision for 2e integrals")')
    rc = load_space('.')
    If (rc /= 0) Call die('input: '//Trim(host_err))
    Write(*, '(4X,"Nd   =",I9,"  Nc =",I7,"  Nst =",I5,"  Nd0 =",I6)') Nd, Nc, Nst, Nd0
    Call write_conf_det('CONF.DET')

    legacy = .false.
    Call Get_Environment_Variable('PCI_CONF_HIJ', env, status=ios)
    If (ios == 0) legacy = (Trim(env) == '1')
    dev = 0
    Call Get_Environment_Variable('PCI_DEVICE', env, status=ios)
    If (ios == 0) Read(env, *) dev
    rc = pci_gpu_create(dev, ctx)
    If (rc /= 0) Call die('pci_gpu_create: '//Trim(pci_error_text()))
    rc = upload(ctx)
    If (rc /= 0) Call die('upload: '//Trim(host_err))
    If (kCSF > 0 .and. Kv == 3) Call die('Kv=3 (J-projection) is not compatible with kCSF>0. '// &
        'CSFs already have definite J; FormJ/Jsq are not built for kCSF runs. Use Kv=4.')

    nbas = Nd
    have_hij = .false.
    ! (a pCONF.HIJ holds H in the determinant basis: a CSF run builds its own)
    If (Kl == 1 .and. kCSF == 0) Inquire(file='pCONF.HIJ', exist=have_hij)
    If (kCSF > 0) Then
        ! jbasis + formh_sym on the GPU (src/pconf.F90:146-149, 211-214)
        Allocate(ndcs(Nconf_neq))
        Call System_Clock(t1)
        rc = pci_gpu_csf_basis(ctx, Nconf_neq, Nc_neq, ndcs, ncsf, nccj)
        If (rc /= 0) Call die('pci_gpu_csf_basis: '//Trim(pci_error_text()))
        Call write_csf_log(ndcs)
        Write(*, '(4x,a,i7,2x,a,i8)') 'Nconf=', Nc, 'Ncsf=', ncsf
        Write(11, '(4x,a,i7,2x,a,i8)') 'Nconf=', Nc, 'Ncsf=', ncsf
        If (ncsf == 0) Then
            Write(*, '(/4x,a,f4.1,a)') 'Term ', Jm, ' is absent in all configurations'
            Write(11, '(/4x,a,f4.1,a)') 'Term ', Jm, ' is absent in all configurations'
            Stop
        End If
        Call timing('CSF basis construction (GPU)', t1)
        Call calc_nd0(ndcs)
        If (Nd0 > 0 .and. Nlv > Nd0) Nlv = Nd0
        nbas = ncsf
        Call System_Clock(t1)
        rc = pci_gpu_build_h_csf(ctx, numh, numzero, hmin)
        If (rc /= 0) Call die('pci_gpu_build_h_csf: '//Trim(pci_error_text()))
        Write(*, '(4X,"NumH = ",I0,/4X,"numzero = ",I0,/4X,"Hmin = ",F14.8)') numh, numzero, hmin
        Write(11, '(4X,"NumH = ",I0,/4X,"numzero = ",I0,/4X,"Hmin = ",F14.8)') numh, numzero, hmin
        Call timing('FormH_sym (GPU)', t1)
    Else If (have_hij) Then
        rc = read_hij('pCONF.HIJ', rowptr, col, val)
        If (rc /= 0) Call die(Trim(host_err))
        numh = rowptr(Nd)
        rc = pci_gpu_set_csr(ctx, Nd, numh, rowptr, col, val)
        If (rc /= 0) Call die('pci_gpu_set_csr: '//Trim(pci_error_text()))
        Write(*, '(4X,"H read from pCONF.HIJ: NumH = ",I0)') numh
        If (legacy) Call write_hij_triples('CONF.HIJ', rowptr, col, val)
    Else
        Call System_Clock(t1)
        rc = pci_gpu_build_h(ctx, 0, numh, numzero, hmin)
        If (rc /= 0) Call die('pci_gpu_build_h: '//Trim(pci_error_text()))
        Write(*, '(4X,"NumH = ",I0,/4X,"numzero = ",I0,/4X,"Hmin = ",F14.8)') numh, numzero, hmin
        Write(11, '(4X,"NumH = ",I0,/4X,"numzero = ",I0,/4X,"Hmin = ",F14.8)') numh, numzero, hmin
        Call timing('FormH (GPU)', t1)
        If (Kw == 1 .or. legacy) Then
            Allocate(rowptr(0:Nd), col(numh), val(numh))
            rc = pci_gpu_get_csr(ctx, rowptr, col, val)
            If (rc /= 0) Call die('pci_gpu_get_csr: '//Trim(pci_error_text()))
            If (Kw == 1) Call write_hij('pCONF.HIJ', rowptr, col, val)
            If (legacy) Call write_hij_triples('CONF.HIJ', rowptr, col, val)
            Deallocate(rowptr, col, val)
        End If
    End If
    If (kCSF == 0) Then
        Call System_Clock(t1)
        rc = pci_gpu_build_j(ctx, numj)
        If (rc /= 0) Call die('pci_gpu_build_j: '//Trim(pci_error_text()))
        If (Kw == 1 .or. legacy) Then
            Allocate(rowptr(0:Nd), col(numj), val(numj))
            rc = pci_gpu_get_jcsr(ctx, rowptr, col, val)
            If (rc /= 0) Call die('pci_gpu_get_jcsr: '//Trim(pci_error_text()))
            If (Kw == 1) Call write_hij('pCONF.JJJ', rowptr, col, val)
            If (legacy) Call write_hij_triples('CONF.JJJ', rowptr, col, val)
            Deallocate(rowptr, col, val)
        End If
        Call timing('FormJ (GPU)', t1)
    End If

    flags = 0
    Call Get_Environment_Variable('PCI_DVD_FLAGS', env, status=ios)
    If (ios == 0) Read(env, *) flags
    If (Kv == 3) Then
        ! selection of states with J = XJ_av (K_prj = 1, src/pconf.F90:371-379)
        njd = jterms(jt, 256)
        If (njd <= 0) Call die('J terms of the space could not be listed')
        rc = pci_gpu_set_jterms(ctx, XJ_av, njd, jt)
        If (rc /= 0) Call die('pci_gpu_set_jterms: '//Trim(pci_error_text()))
        flags = Ior(flags, PCI_DVD_JSEL)
        Write(*, '(4X,"Selection of states with J =",F5.1)') XJ_av
    End If
    If (Iand(flags, PCI_DVD_LEARNED) /= 0) Then
        Call Get_Environment_Variable('PCI_MLP_WEIGHTS', wfile, status=ios)
        If (ios /= 0) Call die('PCI_DVD_LEARNED needs PCI_MLP_WEIGHTS (network file)')
        rc = pci_gpu_mlp_nparams(nw)
        Allocate(w(nw))
        Open(unit=12, file=Trim(wfile), access='stream', form='unformatted', status='old', iostat=ios)
        If (ios /= 0) Call die('cannot open '//Trim(wfile))
        Read(12, iostat=ios) w
        Close(12)
        If (ios /= 0) Call die('short network file '//Trim(wfile))
        rc = pci_gpu_mlp_set_weights(ctx, w, nw)
        If (rc /= 0) Call die('pci_gpu_mlp_set_weights: '//Trim(pci_error_text()))
        Deallocate(w)
    End If
    ! |Kl4| = 2: start from the vectors of an earlier run (FormB0,
    ! src/davidson.f90:136-154)
    ptk0 = c_null_ptr
    px0 = c_null_ptr
    If (Abs(Kl4) == 2) Then
        If (kCSF > 0) Then
            rc = read_start_vectors('CONF.WFS', nbas, Nlv, tk0, x0, nread, ndpt)
        Else
            rc = read_start_vectors('CONF.XIJ', nbas, Nlv, tk0, x0, nread, ndpt)
        End If
        If (rc /= 0) Call die(Trim(host_err))
        Write(*, '(A,I0)') ' Vector length = ', ndpt
        If (nread == 0) Call die('Kl4 = 2: no start vectors read')
        Nlv = nread
        ptk0 = c_loc(tk0)
        px0 = c_loc(x0)
    End If
    prm%nlv = Nlv
    prm%nd0 = Nd0
    prm%crt4 = Crt4
    prm%n_it = N_it
    prm%flags = flags
    prm%refine_every = 0
    prm%reserved = 0
    Allocate(tk(Nlv), tj(Nlv), vecs(nbas, Nlv))
    Call System_Clock(t1)
    rc = pci_gpu_davidson_ex(ctx, prm, ptk0, px0, Max(KXIJ, 0), c_funloc(xij_checkpoint), c_null_ptr, &
                             tk, tj, vecs, st)
    If (rc /= 0) Call die('pci_gpu_davidson: '//Trim(pci_error_text()))
    Write(*, '(4X,"Davidson: ",I0," iterations, max residual ",ES10.3)') st%iterations, st%max_residual
    Call timing('Davidson procedure (GPU)', t1)
    If (kCSF > 0) Then
        ! CSFs have J = Jm (src/pconf.F90:2585-2589); the CSF-basis vectors go
        ! to CONF.WFS, unsym's determinant-basis ones to CONF.XIJ
        ! (src/csf.f90:1088-1134, src/pconf.F90:247-255)
        tj = Jm
        Call write_xij('CONF.WFS', tk, tj, vecs)
        Allocate(dvecs(Nd, Nlv))
        rc = pci_gpu_csf_unsym(ctx, Nlv, vecs, dvecs)
        If (rc /= 0) Call die('pci_gpu_csf_unsym: '//Trim(pci_error_text()))
        Call write_xij('CONF.XIJ', tk, tj, dvecs)
        Call print_energies(6, tk, tj, ncsf)
        Call print_energies(11, tk, tj, ncsf)
    Else
        Call write_xij('CONF.XIJ', tk, tj, vecs)
        Call print_energies(6, tk, tj)
        Call print_energies(11, tk, tj)
    End If
    rc = pci_gpu_destroy(ctx)
    Call timing('Total computation time of pconf_gpu', t0)
    Close(11)

Contains

    Subroutine die(msg)
        Character(Len=*), Intent(In) :: msg
        Write(*, '(A)') ' pconf_gpu: '//Trim(msg)
        Write(11, '(A)') ' pconf_gpu: '//Trim(msg)
        Stop 1
    End Subroutine die

    Subroutine timing(what, ts)
        Character(Len=*), Intent(In) :: what
        Integer(int64), Intent(In) :: ts
        Integer(int64) :: tn
        Call System_Clock(tn)
        Write(*, '(2X,"TIMING >>> ",A," took ",F10.3," s")') what, Real(tn - ts, dp)/Real(rate, dp)
    End Subroutine timing

End Program pconf_gpu
