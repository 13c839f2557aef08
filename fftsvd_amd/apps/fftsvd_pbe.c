/*
 * fftsvd_pbe: the Poisson-Boltzmann solve driven through the drop-in C
 * interface (include/FFTSVDpbeAPI.h), with the reference FFTSVDpbe
 * executable's inputs and outputs (FFTSVDpbe.c:101-268): DelPhi unit files
 * in the working directory
 *    fort.10 parameters, fort.11 radii (SIZ), fort.12 charges (CRG),
 *    fort.13 input PDB, fort.15 collect PDB, fort.16 output FRC,
 *    fort.19 output MODPDB,
 * a flat surface written to operator.srf, the solvation energy on stdout.
 * Plain C: it links only against libfftsvd_hip.so.
 */
#include <stdio.h>
#include <stdlib.h>

#include "FFTSVDpbeAPI.h"

int main(void) {
   unsigned int natoms = 0, ncollect = 0, nsiz = 0, ncrg = 0, a;
   PDBentry *atoms = NULL, *collect = NULL;
   SIZentry* siz = NULL;
   CRGentry* crg = NULL;
   Panel **salt = NULL, **diel = NULL, **dcav = NULL, **scav = NULL;
   unsigned int *nsaltp = NULL, *ndielp = NULL, *ndcavp = NULL, *nscavp = NULL;
   unsigned int nsalt = 0, ndiel = 0, ndcav = 0, nscav = 0, ntotal = 0;
   unsigned int *dielpar = NULL, *dcavpar = NULL, *scavpar = NULL;
   SurfaceOperator op = NULL;
   QualocationOperator qop = NULL;
   Preconditioner pc = NULL;
   Vector rhs = NULL, sol = NULL;
   double energy = 0.0;

   setvbuf(stdout, NULL, _IOLBF, 0);
   printf("fftsvd_pbe: FFT-SVD boundary-element Poisson-Boltzmann solve (MI355X)\n");

   readParams(delphiparameterfilename);
   checkParams();
   readPDB(delphipdbinputfilename, &natoms, &atoms);
   readSIZ(delphisizfilename, &nsiz, &siz);
   readCRG(delphicrgfilename, &ncrg, &crg);
   printf("%u atoms, %u radius and %u charge entries\n", natoms, nsiz, ncrg);
   assignRadiiCharges(atoms, natoms, siz, nsiz, crg, ncrg);
   if (outputmodpdb) writeMODPDB(delphipdboutputfilename, atoms, natoms);

   if (usecurved)
      generateCurvedSRF(delphipdbinputfilename, delphisizfilename, "operator.srf");
   else
      generateFlatSRF(delphipdbinputfilename, delphisizfilename, "operator.srf");
   readSRF("operator.srf", &salt, &nsaltp, &nsalt, &diel, &ndielp, &ndiel, &dielpar, &dcav, &ndcavp, &ndcav, &dcavpar,
           &scav, &nscavp, &nscav, &scavpar, &ntotal);
   printf("%u panels\n", ntotal);

   /* the qualocation or the surface (first-kind) formulation, as FFTSVDpbe.c:169-233 */
   if (usequalocation) {
      printf("qualocation operator\n");
      generateQualocationOperator(&qop, atoms, natoms, diel, ndielp, ndiel, dcav, ndcavp, ndcav, dcavpar, ntotal);
      generateQualocationOperatorPreconditioner(&pc, qop, ntotal);
      generateQualocationRHS(&rhs, qop, ntotal);
      solveQualocationOperator(&sol, qop, pc, rhs, ntotal);
   } else {
      generateSurfaceOperator(&op, atoms, natoms, salt, nsaltp, nsalt, diel, ndielp, ndiel, dielpar, dcav, ndcavp,
                              ndcav, dcavpar, scav, nscavp, nscav, scavpar, ntotal);
      generateSurfaceOperatorPreconditioner(&pc, op, ntotal);
      generateRHS(&rhs, op, ntotal);
      solveSurfaceOperator(&sol, op, pc, rhs, ntotal);
   }

   readPDB(delphipdbcollectfilename, &ncollect, &collect);
   assignRadiiCharges(collect, ncollect, siz, nsiz, crg, ncrg);
   if (usequalocation)
      collectQualocationPotentials(collect, ncollect, atoms, natoms, qop, sol);
   else
      collectPotentials(collect, ncollect, atoms, natoms, op, sol);
   if (outputfrc) writeFRC(delphifrcfilename, collect, ncollect);

   /* reaction-field energy, 1/2 sum q phi, in kcal/mol (potentials are kT/e) */
   for (a = 0; a < ncollect; a++) energy += 0.5 * collect[a].charge * collect[a].potential;
   printf("solvation energy (reaction field, kT/e potentials x 332.112/561): %.10f kcal/mol\n",
          energy * 332.112 / KT_CONVERSION);

   free(rhs);
   free(sol);
   free(atoms);
   free(collect);
   free(siz);
   free(crg);
   return 0;
}
