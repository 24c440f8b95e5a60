! ISO_C_BINDING interface to libpci_gpu.so (include/pci_gpu.h).
!
! This module is what the reference's Fortran would `Use` to call the
! MI355X path in place of FormH / RedistributeHamCSR / FormJ / Diag4
! (src/pconf.F90:217-238).  Arrays are passed by reference (Fortran
! column-major Iarr(Ne,Nd) is exactly the C determinant-major layout).
Module pci_gpu_api
    Use, Intrinsic :: iso_c_binding
    Implicit None

    Integer(c_int), Parameter :: PCI_OK = 0
    Integer(c_int), Parameter :: PCI_DVD_FP32 = 1, PCI_DVD_LEARNED = 2, PCI_DVD_TRACE = 8, &
                                 PCI_DVD_COUPLED = 16, PCI_DVD_JSEL = 32, &
                                 PCI_DVD_NOADMIX = 64
    Integer(c_int), Parameter :: PCI_BUILD_BLOCKS = 1

    Type, bind(C) :: pci_dvd_params
        Integer(c_int32_t) :: nlv
        Integer(c_int32_t) :: nd0
        Real(c_double)     :: crt4
        Integer(c_int32_t) :: n_it
        Integer(c_int32_t) :: flags
        Integer(c_int32_t) :: refine_every
        Integer(c_int32_t) :: reserved
    End Type pci_dvd_params

    Type, bind(C) :: pci_dvd_stats
        Integer(c_int32_t) :: iterations, converged, spmm_fp32, spmm_fp64, learned_used, learned_rejected
        Real(c_double)     :: max_residual, t_total_ms, t_spmm_ms, t_init_ms, spmm_bytes, t_spmm32_ms
        Real(c_double)     :: resid_hist(512)
    End Type pci_dvd_stats

    Interface
        Integer(c_int) Function pci_gpu_create(device, ctx) bind(C, name="pci_gpu_create")
            Import :: c_int, c_ptr
            Integer(c_int), Value :: device
            Type(c_ptr), Intent(Out) :: ctx
        End Function

        Integer(c_int) Function pci_gpu_destroy(ctx) bind(C, name="pci_gpu_destroy")
            Import :: c_int, c_ptr
            Type(c_ptr), Value :: ctx
        End Function

        Type(c_ptr) Function pci_gpu_last_error() bind(C, name="pci_gpu_last_error")
            Import :: c_ptr
        End Function

        Integer(c_int) Function pci_gpu_set_space(ctx, ns, nso, ll, jj, kk, npos, jz, nh, ne, nd, iarr, nc, ndc, &
                                                  mj, gj) bind(C, name="pci_gpu_set_space")
            Import :: c_int, c_int32_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: ns, nso, npos, ne, nd, nc, mj
            Integer(c_int32_t), Intent(In) :: ll(*), jj(*), kk(*), jz(*), nh(*), iarr(*), ndc(*)
            Real(c_double), Value :: gj
        End Function

        Integer(c_int) Function pci_gpu_set_integrals(ctx, nx, nhint, rint1, iint1, ngint, kbrt, rint2, iint2, &
                                                      iint3, nlx, ngaunt, in, gnt) bind(C, name="pci_gpu_set_integrals")
            Import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: nx, nhint, kbrt, nlx, ngaunt
            Integer(c_int64_t), Value :: ngint
            Real(c_double), Intent(In) :: rint1(*), rint2(*), gnt(*)
            Integer(c_int32_t), Intent(In) :: iint1(*), iint2(*), iint3(*), in(*)
        End Function

        Integer(c_int) Function pci_gpu_build_h(ctx, flags, numh, numzero, hmin) bind(C, name="pci_gpu_build_h")
            Import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: flags
            Integer(c_int64_t), Intent(Out) :: numh, numzero
            Real(c_double), Intent(Out) :: hmin
        End Function

        Integer(c_int) Function pci_gpu_get_csr(ctx, rowptr, col, val) bind(C, name="pci_gpu_get_csr")
            Import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int64_t), Intent(Out) :: rowptr(*)
            Integer(c_int32_t), Intent(Out) :: col(*)
            Real(c_double), Intent(Out) :: val(*)
        End Function

        ! one rank's rows of WriteMatrixCSR (col/val: c_null_ptr to size first)
        Integer(c_int) Function pci_gpu_get_csr_rows(ctx, row1, nrows, rowptr, col, val, cap) &
                bind(C, name="pci_gpu_get_csr_rows")
            Import :: c_int, c_int32_t, c_int64_t, c_ptr
            Type(c_ptr), Value :: ctx, col, val
            Integer(c_int32_t), Value :: row1, nrows
            Integer(c_int64_t), Intent(Out) :: rowptr(*)
            Integer(c_int64_t), Value :: cap
        End Function

        Integer(c_int) Function pci_gpu_set_csr(ctx, nd, numh, rowptr, col, val) bind(C, name="pci_gpu_set_csr")
            Import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: nd
            Integer(c_int64_t), Value :: numh
            Integer(c_int64_t), Intent(In) :: rowptr(*)
            Integer(c_int32_t), Intent(In) :: col(*)
            Real(c_double), Intent(In) :: val(*)
        End Function

        Integer(c_int) Function pci_gpu_build_j(ctx, numj) bind(C, name="pci_gpu_build_j")
            Import :: c_int, c_int64_t, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int64_t), Intent(Out) :: numj
        End Function

        Integer(c_int) Function pci_gpu_get_jcsr(ctx, rowptr, col, val) bind(C, name="pci_gpu_get_jcsr")
            Import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int64_t), Intent(Out) :: rowptr(*)
            Integer(c_int32_t), Intent(Out) :: col(*)
            Real(c_double), Intent(Out) :: val(*)
        End Function

        Integer(c_int) Function pci_gpu_davidson(ctx, p, tk, tj, vecs, stats) bind(C, name="pci_gpu_davidson")
            Import :: c_int, c_double, c_ptr, pci_dvd_params, pci_dvd_stats
            Type(c_ptr), Value :: ctx
            Type(pci_dvd_params), Intent(In) :: p
            Real(c_double), Intent(Out) :: tk(*), tj(*), vecs(*)
            Type(pci_dvd_stats), Intent(Out) :: stats
        End Function

        ! Diag4 with the |Kl4| = 2 restart (tk0/x0: NULL or the start energies and
        ! vectors) and the KXIJ checkpoint callback (src/davidson.f90:136-154,
        ! src/pconf.F90:2186-2211)
        Integer(c_int) Function pci_gpu_davidson_ex(ctx, p, tk0, x0, every, cb, user, tk, tj, vecs, stats) &
                bind(C, name="pci_gpu_davidson_ex")
            Import :: c_int, c_int32_t, c_double, c_ptr, c_funptr, pci_dvd_params, pci_dvd_stats
            Type(c_ptr), Value :: ctx
            Type(pci_dvd_params), Intent(In) :: p
            Type(c_ptr), Value :: tk0, x0
            Integer(c_int32_t), Value :: every
            Type(c_funptr), Value :: cb
            Type(c_ptr), Value :: user
            Real(c_double), Intent(Out) :: tk(*), tj(*), vecs(*)
            Type(pci_dvd_stats), Intent(Out) :: stats
        End Function

        Integer(c_int) Function pci_gpu_set_sigma(ctx, ksig, kdsig, nmaxs, lmaxs, nhints, iint1s, rsig, dsig, &
                                                  esig, eps, diag) bind(C, name="pci_gpu_set_sigma")
            Import :: c_int, c_int32_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: ksig, kdsig, nmaxs, lmaxs, nhints
            Integer(c_int32_t), Intent(In) :: iint1s(*)
            Real(c_double), Intent(In) :: rsig(*), dsig(*), esig(*), eps(*)
            Type(c_ptr), Value :: diag
        End Function

        ! RintS for Ksig = 2 (src/pconf.F90:764-840): screening coefficients, SCRC.CON
        Integer(c_int) Function pci_gpu_set_sigma2(ctx, ksym, kmax, nsum, scr, ngints, iint2s, iint3s, rsig2, &
                                                   dsig2, esig2) bind(C, name="pci_gpu_set_sigma2")
            Import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: ksym, kmax, nsum
            Real(c_double), Intent(In) :: scr(*)
            Integer(c_int64_t), Value :: ngints
            Integer(c_int32_t), Intent(In) :: iint2s(*), iint3s(*)
            Real(c_double), Intent(In) :: rsig2(*), dsig2(*), esig2(*)
        End Function

        ! CSF basis, kCSF = 1: jbasis (src/csf.f90:17-407), formh_sym (538-964),
        ! unsym (1088-1134)
        Integer(c_int) Function pci_gpu_csf_basis(ctx, nconf_neq, nc_neq, ndcs, ncsf, nccj) &
                bind(C, name="pci_gpu_csf_basis")
            Import :: c_int, c_int32_t, c_int64_t, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: nconf_neq
            Integer(c_int32_t), Intent(In) :: nc_neq(*)
            Integer(c_int32_t), Intent(Out) :: ndcs(*)
            Integer(c_int32_t), Intent(Out) :: ncsf
            Integer(c_int64_t), Intent(Out) :: nccj
        End Function

        Integer(c_int) Function pci_gpu_build_h_csf(ctx, numh, numzero, hmin) bind(C, name="pci_gpu_build_h_csf")
            Import :: c_int, c_int64_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int64_t), Intent(Out) :: numh, numzero
            Real(c_double), Intent(Out) :: hmin
        End Function

        Integer(c_int) Function pci_gpu_csf_unsym(ctx, nvec, x, y) bind(C, name="pci_gpu_csf_unsym")
            Import :: c_int, c_int32_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: nvec
            Real(c_double), Intent(In) :: x(*)
            Real(c_double), Intent(Out) :: y(*)
        End Function

        Integer(c_int) Function pci_gpu_set_jterms(ctx, xj_av, njd, jt) bind(C, name="pci_gpu_set_jterms")
            Import :: c_int, c_int32_t, c_double, c_ptr
            Type(c_ptr), Value :: ctx
            Real(c_double), Value :: xj_av
            Integer(c_int32_t), Value :: njd
            Integer(c_int32_t), Intent(In) :: jt(*)
        End Function

        Integer(c_int) Function pci_gpu_spmm_variant(ctx, nvec, precision, name, len) &
                bind(C, name="pci_gpu_spmm_variant")
            Import :: c_int, c_int32_t, c_char, c_ptr
            Type(c_ptr), Value :: ctx
            Integer(c_int32_t), Value :: nvec, precision, len
            Character(Kind=c_char), Intent(Out) :: name(*)
        End Function

        Integer(c_int) Function pci_gpu_mlp_nparams(n) bind(C, name="pci_gpu_mlp_nparams")
            Import :: c_int, c_int64_t
            Integer(c_int64_t), Intent(Out) :: n
        End Function

        Integer(c_int) Function pci_gpu_mlp_set_weights(ctx, w, n) bind(C, name="pci_gpu_mlp_set_weights")
            Import :: c_int, c_int64_t, c_float, c_ptr
            Type(c_ptr), Value :: ctx
            Real(c_float), Intent(In) :: w(*)
            Integer(c_int64_t), Value :: n
        End Function

        Integer(c_int) Function pci_gpu_mlp_get_weights(ctx, w, n) bind(C, name="pci_gpu_mlp_get_weights")
            Import :: c_int, c_int64_t, c_float, c_ptr
            Type(c_ptr), Value :: ctx
            Real(c_float), Intent(Out) :: w(*)
            Integer(c_int64_t), Value :: n
        End Function
    End Interface

Contains

    ! Text of the last error as a Fortran string.
    Function pci_error_text() Result(msg)
        Character(Len=512) :: msg
        Type(c_ptr) :: p
        Character(Kind=c_char), Pointer :: s(:)
        Integer :: i
        msg = ' '
        p = pci_gpu_last_error()
        If (.not. c_associated(p)) Return
        Call c_f_pointer(p, s, [512])
        Do i = 1, 512
            If (s(i) == c_null_char) Exit
            msg(i:i) = s(i)
        End Do
    End Function pci_error_text

End Module pci_gpu_api
