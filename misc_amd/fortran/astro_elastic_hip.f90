! astro_elastic_hip.f90 -- GPU element assembly for a Fortran host.
!
! gpuGetElasticSolnLoad(buildMat) has the effect of the reference's
! getElasticSolnLoad (Working/AStrO_commandFunctions.f90:260-351) on the
! reference's own arrays (module AStrO_globalData): with buildMat = 1 the
! element tangents are assembled into elasticMat with elMatCols, and in any
! case elasticLoad(k) is decreased by the internal force R(u) of the current
! nodeDisp (R = K u, the linear static tangent at rest).  The work runs in
! misc_amd/libastro_hip.so (include/astro_abi.h) through ISO_C_BINDING: the
! first call builds a device context from the host's model arrays --
! nodeList, elementList, elementType, the sections, layups and materials
! with the design-variable adjustments of r_getMaterialStiffness,
! r_getMaterialDensity and r_getElThickOffset / r_getShellStiffness
! (Working/AStrO_r_designPropertyFunctions.f90:8-103, 223-330), and the
! constraint terms with their node or node-set names -- and later calls reuse
! it until the design vector r_dVec changes.
!
! Scope: brick8 (8) and shell4 (41) meshes of one element type; isotropic or
! orthotropic engineering constants (a material given by its 21-entry
! stiffness matrix stops the run).  For shell4 the condensed element matrix
! and the nodal load are produced; the internal-mode arrays intElasticMat and
! intElasticLoad are left as they are (zero for a linear step from rest).
! A device or argument error stops the run with the library's message.
module AStrO_gpuElastic
    use, intrinsic :: iso_c_binding
    use AStrO_constantVals
    use AStrO_globalData
    implicit none
    private
    public :: gpuGetElasticSolnLoad, gpuElasticRelease

    type, bind(C) :: astro_model_t
        integer(c_int32_t) :: numNodes, numEls
        type(c_ptr) :: nodeList, elementList, elementType, eprops, density, cond
        integer(c_int32_t) :: numMpcTerms
        type(c_ptr) :: mpcEqn, mpcDof, mpcCoef, mpcIsSet, mpcNodeStart, mpcNodes
        integer(c_int32_t) :: numMPC
        type(c_ptr) :: mpcRHS
        integer(c_int32_t) :: numLoads
        type(c_ptr) :: loadNode, loadVal
        integer(c_int32_t) :: numLayers
        type(c_ptr) :: layers, zOffset, orient
        type(c_ptr) :: specHeat
    end type

    type, bind(C) :: astro_options_t
        integer(c_int32_t) :: maxBW, literalMaxBW, thermLiteralD8, numVecs
        real(c_double) :: rtol
        integer(c_int32_t) :: maxCycles, coupledProfile, precision, stepGraphs
    end type

    interface
        integer(c_int) function astro_ctx_create(model, opts, ctx) bind(C, name='astro_ctx_create')
            import :: c_int, c_ptr, astro_model_t, astro_options_t
            type(astro_model_t), intent(in) :: model
            type(astro_options_t), intent(in) :: opts
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function astro_ctx_destroy(ctx) bind(C, name='astro_ctx_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function astro_ctx_assemble_elastic(ctx, massFactor) bind(C, name='astro_ctx_assemble_elastic')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: massFactor
        end function
        integer(c_int) function astro_ctx_elastic_soln_load(ctx) bind(C, name='astro_ctx_elastic_soln_load')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function astro_ctx_get_f64(ctx, name, host, count) bind(C, name='astro_ctx_get_f64')
            import :: c_int, c_ptr, c_char, c_double, c_int64_t
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: name(*)
            real(c_double), intent(out) :: host(*)
            integer(c_int64_t), value :: count
        end function
        integer(c_int) function astro_ctx_get_i32(ctx, name, host, count) bind(C, name='astro_ctx_get_i32')
            import :: c_int, c_ptr, c_char, c_int32_t, c_int64_t
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: name(*)
            integer(c_int32_t), intent(out) :: host(*)
            integer(c_int64_t), value :: count
        end function
        integer(c_int) function astro_ctx_set_f64(ctx, name, host, count) bind(C, name='astro_ctx_set_f64')
            import :: c_int, c_ptr, c_char, c_double, c_int64_t
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: name(*)
            real(c_double), intent(in) :: host(*)
            integer(c_int64_t), value :: count
        end function
        type(c_ptr) function astro_last_error() bind(C, name='astro_last_error')
            import :: c_ptr
        end function
        integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
            import :: c_size_t, c_ptr
            type(c_ptr), value :: s
        end function
    end interface

    type(c_ptr), save :: ctx = c_null_ptr
    logical, save :: assembled = .false.
    real*8, allocatable, save :: dVecAtSetup(:)

contains

    subroutine check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        integer :: n, i
        character(len=512) :: text
        if (rc .eq. 0) return
        n = int(c_strlen(astro_last_error()))
        call c_f_pointer(astro_last_error(), msg, [n])
        text = ' '
        do i = 1, min(n, 512)
            text(i:i) = msg(i)
        enddo
        write(*,'(A,A,A,I0,A,A)') 'gpuElastic: ', what, ' failed (', rc, '): ', trim(text)
        stop 1
    end subroutine check

    ! Node list of a constraint term: a node number, or the nodes of a node set.
    subroutine term_nodes(name, nodes, cnt, isSet)
        character(len=*), intent(in) :: name
        integer(c_int32_t), intent(out) :: nodes(:)
        integer, intent(out) :: cnt, isSet
        integer :: nd, ios, q, k
        read(name, *, iostat=ios) nd
        if (ios .eq. 0) then
            nodes(1) = nd
            cnt = 1
            isSet = 0
            return
        endif
        isSet = 1
        cnt = 0
        do q = 1, numNdSets
            if (ndSetName(q) .eq. name) then
                do k = ndSetRange(q-1) + 1, ndSetRange(q)
                    cnt = cnt + 1
                    nodes(cnt) = nodeSets(k)
                enddo
                return
            endif
        enddo
        write(*,'(A,A)') 'gpuElastic: unknown node or node set in a constraint: ', trim(name)
        stop 1
    end subroutine term_nodes

    ! The model arrays in the library's layout and the device context.
    subroutine setup()
        type(astro_model_t) :: m
        type(astro_options_t) :: o
        real(c_double), allocatable, target :: xyz(:), ep(:), den(:), cnd(:), lay(:), zo(:), ori(:), coef(:), rhs(:)
        real(c_double), allocatable, target :: sh(:)
        integer(c_int32_t), allocatable, target :: conn(:), et(:), eqn(:), dof(:), isSet(:), nstart(:), nodes(:)
        integer(c_int32_t), allocatable :: tn(:)
        integer :: e, sec, mat, i, j, l, d, nL, nLay, cnt, iset, t, nTerms, k
        real*8 :: lp(6)

        if (c_associated(ctx)) then
            if (numDVar .eq. 0) return
            if (all(r_dVec .eq. dVecAtSetup)) return
            call gpuElasticRelease()
        endif
        allocate(tn(max(1, numNodes)))
        allocate(xyz(3*numNodes), conn(8*numEls), et(numEls), ep(9*numEls), den(numEls), cnd(6*numEls), sh(numEls))
        xyz = reshape(nodeList(1:3, 1:numNodes), [3*numNodes])
        conn = reshape(elementList(1:8, 1:numEls), [8*numEls])
        et = elementType(1:numEls)
        nLay = 0
        do e = 1, numEls
            sec = elementSection(e)
            if (elementType(e) .eq. 41) nLay = max(nLay, secLayupRange(sec) - secLayupRange(sec-1))
        enddo
        allocate(lay(max(1, 6*nLay*numEls)), zo(numEls), ori(6*numEls))
        lay = r_0
        zo = r_0
        ori = r_0
        do e = 1, numEls
            sec = elementSection(e)
            mat = sectionMatId(sec)
            if (elementType(e) .eq. 41) mat = layupMatId(secLayupRange(sec-1) + 1)
            if (abs(materialStiffMat(1, mat)) .gt. 0d0) then
                write(*,'(A)') 'gpuElastic: materials given by a stiffness matrix are not supported'
                stop 1
            endif
            ! engineering constants and density with the element's design
            ! variables (r_getMaterialStiffness, r_getMaterialDensity)
            ep(9*(e-1)+1:9*e) = materialElastic(1:9, mat)
            den(e) = materialDensity(mat)
            cnd(6*(e-1)+1:6*e) = materialThermCond(1:6, mat)
            sh(e) = materialSpecHeat(mat)
            if (numDVar .gt. 0) then
                do i = elToDRange(e-1) + 1, elToDRange(e)
                    d = elToD(i)
                    if (dCategory(d) .eq. 'elastic') then
                        if (dSubCat(d) .eq. 'modulus') then
                            j = dComponent(d)
                        elseif (dSubCat(d) .eq. 'poissonRatio') then
                            j = dComponent(d) + 3
                        else
                            j = dComponent(d) + 6
                        endif
                        ep(9*(e-1)+j) = ep(9*(e-1)+j) + elToCoef(i)*r_dVec(d)
                    elseif (dCategory(d) .eq. 'density') then
                        den(e) = den(e) + elToCoef(i)*r_dVec(d)
                    endif
                enddo
            endif
            if (elementType(e) .eq. 41) then
                ! layers (E1, E2, nu12, G12, thickness, angle) and the z offset
                ! (r_getShellStiffness, r_getElThickOffset)
                nL = 0
                do l = secLayupRange(sec-1) + 1, secLayupRange(sec)
                    nL = nL + 1
                    j = layupMatId(l)
                    lp = (/ materialElastic(1, j), materialElastic(2, j), materialElastic(4, j), &
                            materialElastic(7, j), layupThickness(l), layupAngle(l) /)
                    if (numDVar .gt. 0) then
                        do i = elToDRange(e-1) + 1, elToDRange(e)
                            d = elToD(i)
                            if (dCategory(d) .eq. 'layup' .and. dLayer(d) .eq. nL) then
                                if (dSubCat(d) .eq. 'modulus') then
                                    k = dComponent(d)
                                    lp(k) = lp(k) + elToCoef(i)*r_dVec(d)
                                elseif (dSubCat(d) .eq. 'poissonRatio') then
                                    lp(3) = lp(3) + elToCoef(i)*r_dVec(d)
                                elseif (dSubCat(d) .eq. 'shearModulus') then
                                    lp(4) = lp(4) + elToCoef(i)*r_dVec(d)
                                elseif (dSubCat(d) .eq. 'thickness') then
                                    lp(5) = lp(5) + elToCoef(i)*r_dVec(d)
                                elseif (dSubCat(d) .eq. 'angle') then
                                    lp(6) = lp(6) + elToCoef(i)*r_dVec(d)
                                endif
                            endif
                        enddo
                    endif
                    lay(6*nLay*(e-1) + 6*(nL-1) + 1 : 6*nLay*(e-1) + 6*nL) = lp
                enddo
                zo(e) = sectionZOffset(sec)
                if (numDVar .gt. 0) then
                    do i = elToDRange(e-1) + 1, elToDRange(e)
                        d = elToD(i)
                        if (dCategory(d) .eq. 'layup' .and. dSubCat(d) .eq. 'zOffset') &
                            zo(e) = zo(e) + elToCoef(i)*r_dVec(d)
                    enddo
                endif
                ori(6*(e-1)+1:6*(e-1)+3) = sectionOrient(1:3, 3*sec-2)
                ori(6*(e-1)+4:6*e) = sectionOrient(1:3, 3*sec-1)
            endif
        enddo
        ! constraint terms with their node lists
        nTerms = 0
        if (allocated(mpcEqn)) nTerms = size(mpcEqn)
        allocate(eqn(max(1, nTerms)), dof(max(1, nTerms)), coef(max(1, nTerms)), isSet(max(1, nTerms)), &
                 nstart(nTerms+1), rhs(max(1, numMPC)))
        cnt = 0
        do t = 1, nTerms
            call term_nodes(mpcNode(t), tn, i, iset)
            cnt = cnt + i
        enddo
        allocate(nodes(max(1, cnt)))
        nstart(1) = 0
        do t = 1, nTerms
            call term_nodes(mpcNode(t), tn, i, iset)
            eqn(t) = mpcEqn(t)
            dof(t) = mpcDof(t)
            coef(t) = mpcCoef(t)
            isSet(t) = iset
            nodes(nstart(t)+1:nstart(t)+i) = tn(1:i)
            nstart(t+1) = nstart(t) + i
        enddo
        if (numMPC .gt. 0) rhs(1:numMPC) = mpcRHS(1:numMPC)

        m%numNodes = numNodes
        m%numEls = numEls
        m%nodeList = c_loc(xyz)
        m%elementList = c_loc(conn)
        m%elementType = c_loc(et)
        m%eprops = c_loc(ep)
        m%density = c_loc(den)
        m%cond = c_loc(cnd)
        m%numMpcTerms = nTerms
        m%mpcEqn = c_loc(eqn)
        m%mpcDof = c_loc(dof)
        m%mpcCoef = c_loc(coef)
        m%mpcIsSet = c_loc(isSet)
        m%mpcNodeStart = c_loc(nstart)
        m%mpcNodes = c_loc(nodes)
        m%numMPC = numMPC
        m%mpcRHS = c_loc(rhs)
        m%numLoads = 0
        m%loadNode = c_null_ptr
        m%loadVal = c_null_ptr
        m%specHeat = c_loc(sh)
        m%numLayers = nLay
        if (nLay .gt. 0) then
            m%layers = c_loc(lay)
            m%zOffset = c_loc(zo)
            m%orient = c_loc(ori)
        else
            m%layers = c_null_ptr
            m%zOffset = c_null_ptr
            m%orient = c_null_ptr
        endif
        ! the host's own solverMaxBW and its 32-bit rule (the profile is not
        ! used here; the CSR structure is the same either way)
        o%maxBW = solverMaxBW
        o%literalMaxBW = 1
        o%thermLiteralD8 = 1
        o%numVecs = 0
        o%rtol = 0d0
        o%maxCycles = 0
        o%coupledProfile = 0
        o%precision = 0
        o%stepGraphs = 0
        call check(astro_ctx_create(m, o, ctx), 'astro_ctx_create')
        assembled = .false.
        if (numDVar .gt. 0) then
            if (allocated(dVecAtSetup)) deallocate(dVecAtSetup)
            allocate(dVecAtSetup(numDVar))
            dVecAtSetup = r_dVec
        endif
    end subroutine setup

    subroutine gpuGetElasticSolnLoad(buildMat)
        integer, intent(in) :: buildMat
        real(c_double), allocatable :: ld(:)
        call setup()
        if (buildMat .eq. 1 .or. .not. assembled) then
            call check(astro_ctx_assemble_elastic(ctx, 0d0), 'astro_ctx_assemble_elastic')
            assembled = .true.
        endif
        if (buildMat .eq. 1) then
            call check(astro_ctx_get_f64(ctx, 'elasticMat'//c_null_char, elasticMat, int(elMatSize, c_int64_t)), &
                       'reading elasticMat')
            call check(astro_ctx_get_i32(ctx, 'elMatCols'//c_null_char, elMatCols, int(elMatSize, c_int64_t)), &
                       'reading elMatCols')
        endif
        allocate(ld(elMatDim))
        call check(astro_ctx_set_f64(ctx, 'nodeDisp'//c_null_char, nodeDisp, int(elMatDim, c_int64_t)), &
                   'writing nodeDisp')
        call check(astro_ctx_elastic_soln_load(ctx), 'astro_ctx_elastic_soln_load')
        call check(astro_ctx_get_f64(ctx, 'elasticLoad'//c_null_char, ld, int(elMatDim, c_int64_t)), &
                   'reading elasticLoad')
        elasticLoad(1:elMatDim) = elasticLoad(1:elMatDim) + ld
    end subroutine gpuGetElasticSolnLoad

    ! Frees the device context (a later call builds it again).
    subroutine gpuElasticRelease()
        integer(c_int) :: rc
        if (c_associated(ctx)) rc = astro_ctx_destroy(ctx)
        ctx = c_null_ptr
        assembled = .false.
    end subroutine gpuElasticRelease

end module AStrO_gpuElastic
