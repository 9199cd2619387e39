// polymps_gpu -- the PolyMPS program with every operator on the GPU.
//
// Same start-up, per-step operator sequence and outputs as the reference's
// driver (src/main.cpp:75-216 start-up, :221-469 main loop, :471-556 mesh
// set-up), written against the operator classes of polymps_gpu.hpp.
//
// Usage: polymps_gpu [input.json]
//   Without an argument the program asks for the input name on stdin and
//   reads input/<name>.json, as the reference does.
#include <cstdio>
#include <exception>
#include <vector>

#include "polymps_gpu.hpp"

static std::vector<int> partNearMesh;
static std::vector<double> nodeFRWX, nodeFRWY, nodeFRWZ;   // forced mesh nodes (src/main.cpp:60-62)

// The meshes present, in type order: the reference keeps them in an array
// indexed by meshType (mesh[FORCED] = mesh[2]) that it sizes by their count,
// so a forced mesh without a deformable one falls outside it; here the array
// always has the three entries and only the present ones are used.
static std::vector<int> presentMeshes(const MpsParticleSystem* partSyst) {
    std::vector<int> t{FIXED};
    if (partSyst->femOn) t.push_back(DEFORMABLE);
    if (partSyst->forcedOn) t.push_back(FORCED);
    return t;
}

static void closestPoints(MpsParticleSystem* partSyst, MpsParticle* part, PolygonMesh* mesh) {
    for (int me : presentMeshes(partSyst))
        mesh[me].closestPointPNDBoundaryAABB(partSyst->reS2, partSyst->reL2, part->numParticles, partSyst->weightType,
                                             part->particleType, partSyst->fluid, me, FIXED, DEFORMABLE, FORCED,
                                             part->pos, part->particleAtWallPos, part->mirrorParticlePos,
                                             part->distParticleWall2, part->elementID, part->nearMeshType,
                                             part->polygonNormal);
}

static void initMesh(MpsParticleSystem* partSyst, MpsParticle* part, PolygonMesh* mesh, MpsPndNeigh* pndNeigh,
                     MpsInputOutput* io, MpsBucket* buck) {
    printf("Reading STL File... \n");
    for (int me : presentMeshes(partSyst)) {
        mesh[me].setMeshType(me);
        mesh[me].initPolygonMesh(part->numParticles);
        mesh[me].readPolygonMeshFile(me == FIXED ? io->meshRigidFilename
                                                 : (me == DEFORMABLE ? io->meshDeformableFilename : io->meshForcedFilename));
    }
    printf(" Number of meshes: %d\n", partSyst->numOfMeshs);
    for (int me : presentMeshes(partSyst))
        printf(" after removing duplicates, mesh containts %d vertices and %d faces\n", mesh[me].NV.rows(),
               mesh[me].NF.rows());
    printf("...OK\n");
    for (int me : presentMeshes(partSyst))
        mesh[me].initWijnNeigh((int)partSyst->dim, partSyst->weightType, partSyst->partDist, partSyst->reL,
                               partSyst->reS);
    if (partSyst->forcedOn) {
        const int nn = mesh[FORCED].NV.rows();
        nodeFRWX.resize(nn);
        nodeFRWY.resize(nn);
        nodeFRWZ.resize(nn);
        for (int q = 0; q < nn; q++) {
            nodeFRWX[q] = mesh[FORCED].NV(q, 0);
            nodeFRWY[q] = mesh[FORCED].NV(q, 1);
            nodeFRWZ[q] = mesh[FORCED].NV(q, 2);
        }
        partSyst->uniformVelWall[0] = partSyst->uniformVelWall[1] = partSyst->uniformVelWall[2] = 0.0;
    }
    closestPoints(partSyst, part, mesh);
    partNearMesh.clear();
    mesh[0].updateParticlesNearPolygonMesh(partSyst->reS2, partSyst->reL2, part->numParticles, partSyst->weightType,
                                           part->particleType, partSyst->fluid, part->distParticleWall2,
                                           part->pndWallContribution, part->numNeighWallContribution, partNearMesh,
                                           part->particleNearWall);
    pndNeigh->calcWallNPCD(partSyst, part, buck);
}

static void mainLoop(MpsParticleSystem* partSyst, MpsParticle* part, PolygonMesh* mesh, MpsInputOutput* io,
                     MpsBucket* buck, MpsPndNeigh* partPndNeig, MpsPressure* partPress, MpsShifting* partShift,
                     MpsViscosity* partVisc, MpsParticleCollision* partColl, MpsBoundaryCondition* boundCond,
                     MpsVectorMatrix* vectMatr, MpsParticleVelPos* partVelPos) {
    const bool poly = partSyst->wallType == POLYGON;
    while (true) {
        io->displayInfo(partSyst, part, partPress, 100);
        if (partSyst->numOfIterations % partSyst->iterOutput == 0) {
            io->writeOutputFiles(partSyst, part);
            if (poly) {
                if (partSyst->femOn) mesh[DEFORMABLE].writePolygonMeshFile(DEFORMABLE, io->vtuOutputFoldername,
                                                                         partSyst->fileNumber);
                if (partSyst->forcedOn) mesh[FORCED].writePolygonMeshFile(FORCED, io->vtuOutputFoldername,
                                                                        partSyst->fileNumber);
            }
            partSyst->fileNumber++;
        }
        buck->updateParticlesID(partSyst, part);
        if (partSyst->fluidType == NON_NEWTONIAN) {
            partVisc->calcVolumeFraction(partSyst, part, buck);
            partVisc->calcViscosityInteractionVal(partSyst, part, buck);
        }
        partVisc->calcViscosity(partSyst, part, buck);
        part->addGravity(partSyst, part);
        if (partSyst->relaxPress < 1.0) {
            partPress->predictionPressGradient(partSyst, part, buck);
            if (poly) partPress->predictionWallPressGradient(partSyst, part, buck);
        }
        partVelPos->predictionVelocityPosition(partSyst, part);
        part->checkParticleOutDomain(partSyst);
        if (partSyst->collisionType == PC) partColl->checkParticleCollisions(partSyst, part, buck);
        else partColl->checkDynamicParticleCollisions(partSyst, part, buck);
        if (poly) {
            closestPoints(partSyst, part, mesh);
            partNearMesh.clear();
            mesh[0].updateParticlesNearPolygonMesh(partSyst->reS2, partSyst->reL2, part->numParticles,
                                                   partSyst->weightType, part->particleType, partSyst->fluid,
                                                   part->distParticleWall2, part->pndWallContribution,
                                                   part->numNeighWallContribution, partNearMesh,
                                                   part->particleNearWall);
            partPndNeig->calcWallNPCD(partSyst, part, buck);
        }
        if (partSyst->gradientCorrection) vectMatr->correctionMatrix(partSyst, part, buck);
        partPndNeig->calcPndnNeighNPCD(partSyst, part, buck);
        if (partSyst->pndType == DIFFUSIVE) {
            partPndNeig->calcPndDiffusiveTerm(partSyst, part, buck);
            if (poly) {
                if (partSyst->slipCondition == FREE_SLIP) partPndNeig->calcWallSlipPndDiffusiveTerm(partSyst, part, buck);
                else partPndNeig->calcWallNoSlipPndDiffusiveTerm(partSyst, part, buck);
            }
            partPndNeig->updatePnd(part);
            partPndNeig->meanPndParticlesWallDummySurface(partSyst, part, buck);
        } else if (partSyst->pndType == MEAN_SUM_WIJ) {
            if (poly) partPndNeig->meanWallPnd(partSyst, part, buck);
            partPndNeig->meanPnd(partSyst, part, buck);
        }
        if (partSyst->freeSurfType == PND_ARC) {
            partShift->calcNormalParticles(partSyst, part, buck);
            if (poly) partShift->calcWallNormalParticles(partSyst, part, buck);
        }
        boundCond->updateParticleBC(partSyst, part, buck);
        partPress->calcPress(partSyst, part, buck);
        if (!poly) partPress->extrapolatePressParticlesWallDummy(partSyst, part, buck);
        if (partSyst->gradientCorrection) vectMatr->correctionMatrix(partSyst, part, buck);
        partPress->calcPressGradient(partSyst, part, buck);
        if (poly) partPress->calcWallPressGradient(partSyst, part, buck);
        // src/main.cpp:387-398
        if (partSyst->fluidType == NON_NEWTONIAN) {
            partVisc->calcVolumeFraction(partSyst, part, buck);
            partVisc->calcViscosityInteractionVal(partSyst, part, buck);
            if (poly) {
                if (partSyst->slipCondition == FREE_SLIP)
                    partVisc->calcWallSlipViscosityInteractionVal(partSyst, part, buck);
                else
                    partVisc->calcWallNoSlipViscosityInteractionVal(partSyst, part, buck);
            }
        }
        if (poly) {
            if (partSyst->slipCondition == FREE_SLIP) partVisc->calcWallSlipViscosity(partSyst, part, buck);
            else partVisc->calcWallNoSlipViscosity(partSyst, part, buck);
        }
        if (partSyst->forcedOn)
            mesh[FORCED].updateForcedPolygonMesh(nodeFRWX.data(), nodeFRWY.data(), nodeFRWZ.data(),
                                                 partSyst->uniformVelWall, partSyst->timeStep, partSyst->timeCurrent);
        partVelPos->correctionVelocityPosition(partSyst, part);
        part->checkParticleOutDomain(partSyst);
        if (partSyst->shiftingType == 1) {
            partShift->calcShifting(partSyst, part, buck);
        } else if (partSyst->shiftingType == 2) {
            partShift->calcConcAndConcGradient(partSyst, part, buck);
            if (poly) partShift->calcWallConcAndConcGradient(partSyst, part, buck);
        }
        if (!poly) partVelPos->updateVelocityParticlesWallDummy(partSyst, part, buck);
        part->checkParticleOutDomain(partSyst);
        partSyst->numOfIterations++;
        partSyst->timeCurrent += partSyst->timeStep;
        if (partSyst->timeCurrent > partSyst->timeSimulation) break;
    }
}

int main(int argc, char** argv) {
    printf("Start PolyMPS (MI355X).\n");
    try {
        MpsParticleSystem particleSystem;
        MpsParticle particles;
        MpsInputOutput inputOutput;
        if (argc > 1) inputOutput.jsonPath = argv[1];
        inputOutput.readInputFile(&particleSystem, &particles);
        particles.setParameters(&particleSystem);
        MpsBucket buckets;
        buckets.allocateMemory(&particleSystem, &particles);
        MpsBoundaryCondition boundaryConditions;
        boundaryConditions.setBucketBC(&particleSystem, &particles, &buckets);
        buckets.updateParticlesID(&particleSystem, &particles);
        particles.checkParticleOutDomain(&particleSystem);
        printf("Particle class initialized.\n");

        MpsPndNeigh particlePndNeigh;
        std::vector<PolygonMesh> solidMesh;
        if (particleSystem.wallType == POLYGON) {
            solidMesh.resize(3);   // indexed by meshType
            initMesh(&particleSystem, &particles, solidMesh.data(), &particlePndNeigh, &inputOutput, &buckets);
            printf("Mesh class initialized.\n");
        }
        MpsViscosity particleViscosity;
        MpsParticleCollision particleCollision;
        MpsShifting particleShifting;
        MpsPressure particlePress;
        particlePress.allocateMemory(&particles);
        MpsVectorMatrix vectorMatrix;
        MpsParticleVelPos particleVelPos;

        printf("Initial Step... ");
        particlePndNeigh.setInitialPndNumberOfNeighNPCD(&particleSystem, &particles, &buckets);
        if (particleSystem.wallType == POLYGON) particlePndNeigh.meanWallPnd(&particleSystem, &particles, &buckets);
        particlePndNeigh.meanPnd(&particleSystem, &particles, &buckets);
        particlePndNeigh.meanNeighFluidPnd(&particleSystem, &particles, &buckets);
        if (particleSystem.freeSurfType == PND_ARC) {
            particleShifting.calcNormalParticles(&particleSystem, &particles, &buckets);
            if (particleSystem.wallType == POLYGON)
                particleShifting.calcWallNormalParticles(&particleSystem, &particles, &buckets);
        }
        boundaryConditions.updateParticleBC(&particleSystem, &particles, &buckets);
        particlePress.calcPress(&particleSystem, &particles, &buckets);
        inputOutput.writePvd(&particleSystem, &particles);
        inputOutput.deleteDirectoryFiles();
        inputOutput.writeBuckets(&particleSystem, &particles);
        printf("OK\n");

        printf("\nStart Main Loop of Simulation\n\n");
        mainLoop(&particleSystem, &particles, solidMesh.empty() ? nullptr : solidMesh.data(), &inputOutput, &buckets,
                 &particlePndNeigh, &particlePress, &particleShifting, &particleViscosity, &particleCollision,
                 &boundaryConditions, &vectorMatrix, &particleVelPos);
        inputOutput.finishOutput(&particles);
    } catch (const std::exception& e) {
        fprintf(stderr, "polymps_gpu: %s\n", e.what());
        return 1;
    }
    printf("End PolyMPS.\n");
    return 0;
}
