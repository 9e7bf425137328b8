! ref_driver.f90 -- instrumented driver for the literal reference (test
! infrastructure only; see oracle/Makefile).  It is linked against the
! reference's own modules compiled from source and runs the same sequence as
! the reference's *readModelInput + *solve commands (Working/AStrO_runJob.f90:48-65,
! 101-193 and Working/AStrO_commandFunctions.f90:751-782), dumping every array
! of the hot path as raw little-endian binary so that tests can compare the
! oracle restatement and the HIP library against it.
!
! usage: astro_ref_driver <model.yaml> <out-prefix> <elastic|thermal|prep>
!                         [numVecs maxIt solverMaxBW]
! (prep stops after analysisPrep and its dumps).  The optional integers, when
! positive, replace the GMRES restart length and iteration cap of stepSolve
! (Working/AStrO_commandFunctions.f90:711-715, 732-733) and the solverMaxBW of
! the *solve block, so that a job can stop GMRES after a chosen number of full
! restart cycles.
!
! With ASTRO_DV_FILE set, the design-variable file it names is read after the
! model (readDesignVarInput, Working/AStrO_input.f90:1099-1408); the drop-in
! build then dumps getdRelasticdD's column of every variable at the solution
! (dRudD, elMatDim x numDVar), formed by gpuGetdRelasticdD.
!
! With ASTRO_OBJ_FILE set, the objective file it names is read too
! (readObjectiveInput); with ASTRO_WRITE_PREFIX set, the reference's own
! writers run after the solve on its arrays (writeNodeResults,
! writeElementResults, writeObjective after getObjective, writeDVarValues;
! Working/AStrO_output.f90), into files named by that prefix.
!
! Solver parameters follow the *solve block: solverBlockDim = numNodes and
! solverMaxBW from 27*numNodes*numNodes in default INTEGER arithmetic
! (Working/AStrO_runJob.f90:107-114).  Thermal mode sets thermMatDim =
! numNodes, which the reference never assigns (it is read at
! Working/AStrO_commandFunctions.f90:690, 767), so that its thermal path can be
! exercised at all.
program astro_ref_driver
    use AStrO_constantVals
    use AStrO_globalData
    use AStrO_input
    use AStrO_bookKeeping
    use AStrO_commandFunctions
    use AStrO_solvers
    use AStrO_output
    use AStrO_objective
#ifdef ASTRO_HIP
    use AStrO_gpuElastic
#endif
    implicit none

    character(len=128) :: modelFile
    character(len=256) :: prefix
    character(len=16) :: mode, argv
    character(len=256) :: dvFile, tFile
    integer :: tStat
    character(len=128) :: objFile, wrPrefix, wrName
    character(len=16) :: wf(9)
    integer :: dvStat, kdv, objStat, wrStat, q
    integer, allocatable :: allIds(:)
    real*8, allocatable :: dRAll(:)
    integer :: nnd2, bw3, ticks0, ticks1, rate, tu
    integer :: nvArg, itArg, bwArg
    real*8 :: tPrep, tAsm, tScale, tProf, tFact, tSolve

    call get_command_argument(1, modelFile)
    call get_command_argument(2, prefix)
    call get_command_argument(3, mode)
    nvArg = 0
    itArg = 0
    bwArg = 0
    if (command_argument_count() .ge. 6) then
        call get_command_argument(4, argv)
        read(argv, *) nvArg
        call get_command_argument(5, argv)
        read(argv, *) itArg
        call get_command_argument(6, argv)
        read(argv, *) bwArg
    endif

    lfUnit = 30
    open(unit=lfUnit, file=trim(prefix)//'log.txt', action='write', status='replace')

    call readModelInput(modelFile)
    call get_environment_variable('ASTRO_DV_FILE', dvFile, status=dvStat)
    if (dvStat .eq. 0) call readDesignVarInput(dvFile)
    call get_environment_variable('ASTRO_OBJ_FILE', objFile, status=objStat)
    if (objStat .eq. 0) call readObjectiveInput(objFile)
    call get_environment_variable('ASTRO_WRITE_PREFIX', wrPrefix, status=wrStat)

    solverBlockDim = numNodes
    nnd2 = 27*numNodes*numNodes
    solverMaxBW = 1
    bw3 = 1
    do while(bw3 .lt. nnd2)
        solverMaxBW = solverMaxBW + 1
        bw3 = solverMaxBW*solverMaxBW*solverMaxBW
    enddo
    if (bwArg .gt. 0) solverMaxBW = bwArg
    solveThermal = 0
    solveElastic = 1
    nLGeom = 0
    dynamic = 0
    writeSolnHist = 0
    numTSteps = 0
    nMBeta = 0.25d0
    nMGamma = 0.5d0
    delT = 0d0
    simPeriod = 0d0
    rayCoefK = 0d0
    rayCoefM = 0d0
    loadTime = 0d0
    if (trim(mode) .eq. 'thermal') then
        solveThermal = 1
        solveElastic = 0
    endif

    ! buildElasticMPC / buildThermalMPC assign zero-length sections of their
    ! range array even when they allocated nothing (no constraint of that kind,
    ! Working/AStrO_bookKeeping.f90:525-529, 670-674); gfortran never touches
    ! memory there, flang's runtime rejects the unallocated target.  Give the
    ! kind this job has no constraints of an empty array up front.
    if (trim(mode) .eq. 'thermal') then
        allocate(elMPCMatRange(0:0))
    elseif (trim(mode) .ne. 'thermoelastic') then
        allocate(thermMPCMatRange(0:0))
    endif

    call system_clock(ticks0, rate)
    call analysisPrep()
    call system_clock(ticks1)
    tPrep = dble(ticks1 - ticks0)/dble(rate)
    thermMatDim = numNodes
    ! Same class of defect for models without internal DOFs: loadInitialState
    ! and stepSolve clear prevIntDisp / intElasticLoad
    ! (Working/AStrO_commandFunctions.f90:663, 717), which analysisPrep only
    ! allocates when intVecSize > 0 (Working/AStrO_bookKeeping.f90:1005).
    if (.not. allocated(prevIntDisp)) allocate(prevIntDisp(0))
    if (.not. allocated(intElasticLoad)) allocate(intElasticLoad(0))
    ! Arrays the reference allocates and then reads without ever assigning:
    ! the (0) entry of the constraint range arrays (set nowhere in
    ! Working/AStrO_bookKeeping.f90:516-529, 656-674) and the CSR column arrays,
    ! whose 0 entries mark free slots for the assembly probe
    ! (allocated at :374-375, :438-439, probed at
    ! Working/AStrO_commandFunctions.f90:326, 50).  A fresh process heap
    ! usually returns them zeroed; once another runtime has used the heap it
    ! does not, so the driver zeroes them.
    if (allocated(elMPCMatRange)) elMPCMatRange(0) = 0
    if (allocated(thermMPCMatRange)) thermMPCMatRange(0) = 0
    elMatCols(:) = 0
    thermMatCols(:) = 0

    call dumpi(currentRank, numNodes, 'currentRank')
    call dumpi(nDofIndex, 6*numNodes, 'nDofIndex')
    call dumpi(elMatRange, elMatDim+1, 'elMatRange')
    call dumpi(elMatLTRange, elMatDim+1, 'elMatLTRange')
    call dumpi(thermMatRange, numNodes+1, 'thermMatRange')
    call dumpi(thermMatLTRange, numNodes+1, 'thermMatLTRange')
    call dumpi((/solverMaxBW, elMatDim, elMPCDim, thermMPCDim/), 4, 'scalars')
    if (elMPCDim .gt. 0) then
        call dumpi(elMPCMatRange, elMPCDim+1, 'elMPCMatRange0')
        call dumpi(elMPCMatCols, elMPCSize, 'elMPCMatCols')
        call dumpd(elMPCMat, elMPCSize, 'elMPCMat0')
    endif

    if (trim(mode) .eq. 'prep') then
        close(lfUnit)
        stop
    endif

    call loadInitialState()

    if (trim(mode) .eq. 'nonlinear') then
        ! solve() with nonLinearGeom (Working/AStrO_commandFunctions.f90:770-778):
        ! the tangent at rest scales the penalty rows, then stepSolve's elastic
        ! loop (:704-747) assembles and factors the tangent at u every step
        call getElasticSolnLoad(1)
        call scaleElasticMPC()
        nLGeom = 1
        call nonlinearElasticSteps()
        call dumpd(nodeDisp, elMatDim, 'nodeDisp')
        close(lfUnit)
        stop
    endif

    if (trim(mode) .eq. 'thermoelastic') then
        ! solve()'s elastic sequence (Working/AStrO_commandFunctions.f90:751-780):
        ! the tangent is assembled and factored at the zero nodeTemp the
        ! allocation left; then the nodal temperatures a thermal step would
        ! have produced are read from ASTRO_TEMP_FILE (numNodes doubles by
        ! rank; the reference's own thermal step is not used: D8 drops a
        ! coupling of every row) and stepSolve's elastic step runs, its
        ! residual holding the thermal strain
        call getElasticSolnLoad(1)
        call scaleElasticMPC()
        call convertToLTri(elMatLT,elMatLTRange,elMatLTSize,elasticMat,elMatCols,elMatRange, &
             elMatSize,elMatDim,elMPCMat,elMPCMatCols,elMPCMatRange,elMPCSize,elMPCDim)
        call getSparseLDLFact(elMatLT,elMatLTSize,elMatLTRange,elMatDim)
        call get_environment_variable('ASTRO_TEMP_FILE', tFile, status=tStat)
        if (tStat .ne. 0) then
            write(*,'(A)') 'ref_driver: thermoelastic needs ASTRO_TEMP_FILE'
            stop 1
        endif
        open(newunit=tu, file=trim(tFile), access='stream', form='unformatted', action='read')
        read(tu) nodeTemp(1:numNodes)
        close(tu)
        call elasticStep()
        call dumpd(elasticLoad, elMatDim, 'elasticLoad')
        call dumpd(nodeDisp, elMatDim, 'nodeDisp')
        close(lfUnit)
        stop
    endif

    if (solveElastic .eq. 1) then
#ifdef ASTRO_HIP
        ! drop-in build: the reference's own assembly on the same arrays is
        ! dumped for comparison, then the GPU assembly replaces it
        call getElasticSolnLoad(1)
        call dumpd(elasticMat, elMatSize, 'elasticMat_host')
        call dumpi(elMatCols, elMatSize, 'elMatCols_host')
        elasticMat(:) = r_0
        elMatCols(:) = 0
        call system_clock(ticks0)
        call gpuGetElasticSolnLoad(1)
#else
        call system_clock(ticks0)
        call getElasticSolnLoad(1)
#endif
        call system_clock(ticks1)
        tAsm = dble(ticks1 - ticks0)/dble(rate)
        call dumpd(elasticMat, elMatSize, 'elasticMat')
        call dumpi(elMatCols, elMatSize, 'elMatCols')
        call system_clock(ticks0)
        call scaleElasticMPC()
        call system_clock(ticks1)
        tScale = dble(ticks1 - ticks0)/dble(rate)
        if (elMPCDim .gt. 0) then
            call dumpd(elMPCMat, elMPCSize, 'elMPCMat')
            call dumpd(elMPCRHS, elMPCDim, 'elMPCRHS')
        endif
        call system_clock(ticks0)
        call convertToLTri(elMatLT,elMatLTRange,elMatLTSize,elasticMat,elMatCols,elMatRange, &
             elMatSize,elMatDim,elMPCMat,elMPCMatCols,elMPCMatRange,elMPCSize,elMPCDim)
        call system_clock(ticks1)
        tProf = dble(ticks1 - ticks0)/dble(rate)
        call dumpd(elMatLT, elMatLTSize, 'elMatLT')
        call system_clock(ticks0)
        call getSparseLDLFact(elMatLT,elMatLTSize,elMatLTRange,elMatDim)
        call system_clock(ticks1)
        tFact = dble(ticks1 - ticks0)/dble(rate)
        call dumpd(elMatLT, elMatLTSize, 'elMatLTfact')
        call system_clock(ticks0)
        call elasticStep()
        call system_clock(ticks1)
        tSolve = dble(ticks1 - ticks0)/dble(rate)
        call dumpd(elasticLoad, elMatDim, 'elasticLoad')
        call dumpd(nodeDisp, elMatDim, 'nodeDisp')
#ifdef ASTRO_HIP
        ! the solution load at the converged displacement (the reactionForce
        ! output's call, Working/AStrO_runJob.f90:228-233), host and GPU
        elasticLoad(:) = r_0
        call getElasticSolnLoad(0)
        call dumpd(elasticLoad, elMatDim, 'solnLoad_host')
        elasticLoad(:) = r_0
        call gpuGetElasticSolnLoad(0)
        call dumpd(elasticLoad, elMatDim, 'solnLoad')
        if (numDVar .gt. 0) then
            allocate(dRAll(elMatDim*numDVar))
            do kdv = 1, numDVar
                call gpuGetdRelasticdD(loadTime, kdv)
                dRAll((kdv-1)*elMatDim+1 : kdv*elMatDim) = dRudD(1:elMatDim)
            enddo
            call dumpd(dRAll, elMatDim*numDVar, 'dRudD')
        endif
        call gpuElasticRelease()
#endif
    else
        call system_clock(ticks0)
        call getThermalSolnLoad(1)
        call system_clock(ticks1)
        tAsm = dble(ticks1 - ticks0)/dble(rate)
        call dumpd(thermMat, thermMatSize, 'thermMat')
        call dumpi(thermMatCols, thermMatSize, 'thermMatCols')
        call system_clock(ticks0)
        call scaleThermalMPC()
        call system_clock(ticks1)
        tScale = dble(ticks1 - ticks0)/dble(rate)
        if (thermMPCDim .gt. 0) then
            call dumpd(thermMPCMat, thermMPCSize, 'thermMPCMat')
            call dumpd(thermMPCRHS, thermMPCDim, 'thermMPCRHS')
        endif
        call system_clock(ticks0)
        call convertToLTri(thermMatLT,thermMatLTRange,thermMatLTSize,thermMat,thermMatCols,thermMatRange, &
             thermMatSize,thermMatDim,thermMPCMat,thermMPCMatCols,thermMPCMatRange,thermMPCSize,thermMPCDim)
        call system_clock(ticks1)
        tProf = dble(ticks1 - ticks0)/dble(rate)
        call dumpd(thermMatLT, thermMatLTSize, 'thermMatLT')
        call system_clock(ticks0)
        call getSparseLDLFact(thermMatLT,thermMatLTSize,thermMatLTRange,thermMatDim)
        call system_clock(ticks1)
        tFact = dble(ticks1 - ticks0)/dble(rate)
        call dumpd(thermMatLT, thermMatLTSize, 'thermMatLTfact')
        call system_clock(ticks0)
        call thermalStep()
        call system_clock(ticks1)
        tSolve = dble(ticks1 - ticks0)/dble(rate)
        call dumpd(thermalLoad, numNodes, 'thermalLoad')
        call dumpd(nodeTemp, numNodes, 'nodeTemp')
    endif

    if (wrStat .eq. 0) call writeAll()

    open(newunit=tu, file=trim(prefix)//'times.txt', action='write', status='replace')
    write(tu,'(A,ES14.6)') 'prep ', tPrep
    write(tu,'(A,ES14.6)') 'assemble ', tAsm
    write(tu,'(A,ES14.6)') 'scale ', tScale
    write(tu,'(A,ES14.6)') 'profile ', tProf
    write(tu,'(A,ES14.6)') 'factor ', tFact
    write(tu,'(A,ES14.6)') 'solve ', tSolve
    close(tu)
    close(lfUnit)

contains

    ! Linear-static elastic branch of stepSolve (Working/AStrO_commandFunctions.f90:704-747)
    ! with maxNLIt = 1.  getElasticAppliedLoad is replaced by its nodal-load
    ! branch (:368-392) because its design-variable node-load loop reads through
    ! an uninitialised index (r_getNodeDLoad, Working/AStrO_r_designPropertyFunctions.f90:980),
    ! which faults under flang; with no design variables that loop adds nothing.
    subroutine elasticStep()
        integer :: numVecs, maxIt
        if (elMatDim .lt. 200) then
            numVecs = elMatDim/2
        else
            numVecs = 100
        endif
        maxIt = elMatDim
        if (nvArg .gt. 0) numVecs = nvArg
        if (itArg .gt. 0) maxIt = itArg
        elasticLoad(:) = r_0
        intElasticLoad(:) = r_0
#ifdef ASTRO_HIP
        call gpuGetElasticSolnLoad(nLGeom)
#else
        call getElasticSolnLoad(nLGeom)
#endif
        call nodalLoads()
        call getElasticConstraintLoad()
        call gMRes(delDisp,elasticMat,elMatCols,elMatRange,elMatSize,elMatDim,elMPCMat,elMPCMatCols, &
            elMPCMatRange,elMPCSize,elMPCDim,elMatLT,elMatLTRange,elMatLTSize,elasticLoad,numVecs,maxIt)
        nodeDisp(:) = nodeDisp(:) + delDisp(:)
    end subroutine elasticStep

    ! Elastic branch of stepSolve with nonLinearGeom (Working/AStrO_commandFunctions.f90:
    ! 704-747): at most 20 steps, each assembling the tangent and the residual at
    ! u, factoring, solving and adding the increment, until max|du| <= 1e-12.
    ! Nodal loads as elasticStep; the step count is dumped as 'nlSteps'.
    subroutine nonlinearElasticSteps()
        integer :: numVecs, maxIt, it
        real*8 :: delUNorm
        if (elMatDim .lt. 200) then
            numVecs = elMatDim/2
        else
            numVecs = 100
        endif
        maxIt = elMatDim
        if (nvArg .gt. 0) numVecs = nvArg
        if (itArg .gt. 0) maxIt = itArg
        it = 0
        delUNorm = 1d0
        do while (it .lt. 20 .and. delUNorm .gt. 1d-12)
            elasticLoad(:) = r_0
            intElasticLoad(:) = r_0
            call getElasticSolnLoad(1)
            call nodalLoads()
            call getElasticConstraintLoad()
            call convertToLTri(elMatLT,elMatLTRange,elMatLTSize,elasticMat,elMatCols,elMatRange, &
                 elMatSize,elMatDim,elMPCMat,elMPCMatCols,elMPCMatRange,elMPCSize,elMPCDim)
            call getSparseLDLFact(elMatLT,elMatLTSize,elMatLTRange,elMatDim)
            call gMRes(delDisp,elasticMat,elMatCols,elMatRange,elMatSize,elMatDim,elMPCMat,elMPCMatCols, &
                elMPCMatRange,elMPCSize,elMPCDim,elMatLT,elMatLTRange,elMatLTSize,elasticLoad,numVecs,maxIt)
            nodeDisp(:) = nodeDisp(:) + delDisp(:)
            delUNorm = maxval(abs(delDisp))
            it = it + 1
        enddo
        call dumpi((/it/), 1, 'nlSteps')
    end subroutine nonlinearElasticSteps

    ! Thermal branch of stepSolve (Working/AStrO_commandFunctions.f90:689-703); the
    ! models driven here carry no thermal loads, only temperature constraints.
    subroutine thermalStep()
        integer :: numVecs
        if (thermMatDim .lt. 200) then
            numVecs = thermMatDim/2
        else
            numVecs = 100
        endif
        thermalLoad(:) = r_0
        call getThermalSolnLoad(0)
        call getThermalConstraintLoad()
        call gMRes(delDisp(1:numNodes),thermMat,thermMatCols,thermMatRange,thermMatSize,thermMatDim, &
            thermMPCMat,thermMPCMatCols,thermMPCMatRange,thermMPCSize,thermMPCDim,thermMatLT,thermMatLTRange, &
            thermMatLTSize,thermalLoad,numVecs,thermMatDim)
        nodeTemp(:) = nodeTemp(:) + delDisp(1:numNodes)
    end subroutine thermalStep

    ! Nodal loads applied at loadTime: node number or node-set name per line.
    subroutine nodalLoads()
        integer :: l, t, nd, d, k, q, ios
        do l = 1, numLds
            if (loadTime .ge. loadsActTime(1,l) .and. loadTime .le. loadsActTime(2,l)) then
                if (loadType(l) .eq. 'nodal') then
                    do t = loadsRange(l-1) + 1, loadsRange(l)
                        read(loadNodes(t), *, iostat=ios) nd
                        if (ios .eq. 0) then
                            do d = 1, 6
                                k = nDofIndex(d, nd)
                                if (k .ne. 0) elasticLoad(k) = elasticLoad(k) + inputLoads(d, t)
                            enddo
                        else
                            do q = 1, numNdSets
                                if (ndSetName(q) .eq. loadNodes(t)) then
                                    do k = ndSetRange(q-1) + 1, ndSetRange(q)
                                        nd = nodeSets(k)
                                        do d = 1, 6
                                            if (nDofIndex(d, nd) .ne. 0) elasticLoad(nDofIndex(d, nd)) = &
                                                elasticLoad(nDofIndex(d, nd)) + inputLoads(d, t)
                                        enddo
                                    enddo
                                endif
                            enddo
                        endif
                    enddo
                endif
            endif
        enddo
    end subroutine nodalLoads

    ! The reference's writers on its state after the solve.
    subroutine writeAll()
        allocate(allIds(max(numNodes, numEls)))
        do q = 1, size(allIds)
            allIds(q) = q
        enddo
        if (solveElastic .eq. 1) then
            wf(1) = 'displacement'
            wf(2) = 'reactionForce'
            wrName = trim(wrPrefix)//'nodes.yaml'
            call writeNodeResults(wrName, wf, 2, allIds(1:numNodes), numNodes, 0d0, 0)
            wf(1) = 'stress'
            wf(2) = 'strain'
            wf(3) = 'strainEnergyDen'
            wrName = trim(wrPrefix)//'elements.yaml'
            call writeElementResults(wrName, wf, 3, allIds(1:numEls), numEls, 0d0, 0)
        else
            wf(1) = 'temperature'
            wf(2) = 'reactionHeatGen'
            thermalLoad(:) = r_0
            call getThermalSolnLoad(0)
            call dumpd(thermalLoad, numNodes, 'thermalLoadSoln')
            wrName = trim(wrPrefix)//'nodes.yaml'
            call writeNodeResults(wrName, wf, 2, allIds(1:numNodes), numNodes, 0d0, 0)
            wf(1) = 'heatFlux'
            wf(2) = 'tempGradient'
            wrName = trim(wrPrefix)//'elements.yaml'
            call writeElementResults(wrName, wf, 2, allIds(1:numEls), numEls, 0d0, 0)
        endif
        if (objStat .eq. 0) then
            if (numObjTerms .gt. 0) then
                objVal(:) = r_0
                call getObjective(loadTime)
                wf(1) = 'category'
                wf(2) = 'operator'
                wf(3) = 'component'
                wf(4) = 'layer'
                wf(5) = 'coefficient'
                wf(6) = 'exponent'
                wf(7) = 'elementSet'
                wf(8) = 'nodeSet'
                wrName = trim(wrPrefix)//'objective.yaml'
                call writeObjective(wrName, wf, 8, 0)
                call dumpd(objVal, numObjTerms, 'objVal')
            endif
        endif
        if (numDVar .gt. 0) then
            wf(1) = 'category'
            wf(2) = 'subCategory'
            wf(3) = 'component'
            wf(4) = 'layer'
            wf(5) = 'activeTime'
            wrName = trim(wrPrefix)//'dvars.yaml'
            call writeDVarValues(wrName, wf, 5)
        endif
        ! the binary solution history record of step 3 (writeBinarySolution,
        ! Working/AStrO_output.f90:760-782) for a state with distinct values
        if (solveElastic .eq. 1) then
            do q = 1, numNodes
                prevTemp(q) = 0.5d0*q
                prevTdot(q) = -0.25d0*q
            enddo
            prevDisp(:) = nodeDisp(:)
            prevVel(:) = 2d0*nodeDisp(:)
            prevAcc(:) = -3d0*nodeDisp(:)
            call writeBinarySolution(3)
        endif
    end subroutine writeAll

    subroutine dumpi(a, n, name)
        integer, intent(in) :: n
        integer, intent(in) :: a(n)
        character(len=*), intent(in) :: name
        integer :: u
        open(newunit=u, file=trim(prefix)//name//'.i32', access='stream', form='unformatted', &
             status='replace', action='write')
        write(u) a
        close(u)
    end subroutine dumpi

    subroutine dumpd(a, n, name)
        integer, intent(in) :: n
        real*8, intent(in) :: a(n)
        character(len=*), intent(in) :: name
        integer :: u
        open(newunit=u, file=trim(prefix)//name//'.f64', access='stream', form='unformatted', &
             status='replace', action='write')
        write(u) a
        close(u)
    end subroutine dumpd

end program astro_ref_driver
