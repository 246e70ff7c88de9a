//------------------------------------------------------------------------------
// ref_driver.cpp -- test infrastructure only.
//
// Drives the reference VCP3D_CurviLinear build (compiled from the reference
// sources by oracle/Makefile into oracle/_ref/) through its public API, in the
// same order as the reference main (VCP3D_CurviLinear/src/VCP3D_CurviLinear.cpp:
// 23-178), to write golden data:
//   <SolFile>_Res_0        modal residual R(U_0) after the first Residual call
//                          (SolverClass::WriteRestartSolution(0,true),
//                           SolverClass.cpp:4456-4474)
//   <SolFile>_Sol_<n>      nodal dimensional conservative solution after n steps
//   <SolFile>_Res_<n>      modal residual R(U_n) (first-same-as-last residual)
// and prints one JSON line with the wall time of every
// ExecuteTimeSynchronizationStep call (the CPU baseline).
//
// usage: vcp_ref_driver <param.cfg> <nSteps> [dumpStep,dumpStep,...]
//------------------------------------------------------------------------------
#include "VCP3D_CurviLinear.hpp"
#include <cstdio>
#include <sstream>

// Global variables of the reference (defined next to its main()).
int rank, nRanks;
ENUM_SIGNAL CurrentSignal;
su2double rhoDim;
su2double pDim;
su2double pRef;
su2double uRef;
su2double rhoRef;
su2double mu;

static void renameTo(const std::string &solFile, int step, const char *tag)
{
  std::ostringstream src, dst;
  src << solFile << "_TimeStep_" << step;
  dst << solFile << "_" << tag << "_" << step;
  if(std::rename(src.str().c_str(), dst.str().c_str()) != 0)
  {
    std::fprintf(stderr, "rename %s failed\n", src.str().c_str());
    std::exit(2);
  }
}

int main(int argc, char **argv)
{
  if(argc < 3)
  {
    std::fprintf(stderr, "usage: %s param.cfg nSteps [dumpSteps]\n", argv[0]);
    return 1;
  }
  rank = 0; nRanks = 1;
  CurrentSignal = NoSignal;

  const int nSteps = std::atoi(argv[2]);
  std::set<int> dumps;
  if(argc > 3)
  {
    std::stringstream ss(argv[3]);
    std::string tok;
    while(std::getline(ss, tok, ',')) if(tok.size()) dumps.insert(std::atoi(tok.c_str()));
  }

  InputParamClass inputParam;
  inputParam.ReadInputParameters(argv[1]);
  if( inputParam.mSpongeLayer || inputParam.mCharacteristicMatchingLayer )
    inputParam.ReadSpongeLayerData(argv[1]);
  inputParam.ReadSubfaceInformation(argv[1]);
  inputParam.ReadBoundaryData(argv[1]);
  inputParam.ComputeReferenceConditions();

  SolverClass solver(&inputParam);
  solver.ReadGrid();
  solver.ComputeMetricTerms();
  solver.BoundingBoxFluctuations();
  solver.InitSol();
  if( inputParam.mRestart ) solver.ReadRestartSolution();
  if( inputParam.mSpongeLayer ) solver.InitSpongeLayer();
  solver.PrescribedDataIntegrationPoints();
  solver.AddFluctuations();
  if( inputParam.mNSCBC_Specified ) solver.InitNSCBC();
  if( inputParam.mNSCBC_Specified ) solver.AverageDataNSCBC();

  su2double monitoringData[8];
  solver.Residual(monitoringData, true);
  solver.WriteRestartSolution(0, true);
  renameTo(inputParam.mSolFile, 0, "Res");

  std::vector<double> times;
  for(int step=1; step<=nSteps; ++step)
  {
    const double t0 = GetWallTime();
    solver.ExecuteTimeSynchronizationStep(monitoringData);
    times.push_back(GetWallTime() - t0);
    if( dumps.count(step) )
    {
      solver.WriteRestartSolution(step);
      renameTo(inputParam.mSolFile, step, "Sol");
      solver.WriteRestartSolution(step, true);
      renameTo(inputParam.mSolFile, step, "Res");
    }
  }

  std::printf("{\"ref_driver\": 1, \"steps\": %d, \"uRef\": %.17g, \"pRef\": %.17g, "
              "\"rhoRef\": %.17g, \"mu\": %.17g, \"monitoring\": [", nSteps,
              (double) uRef, (double) pRef, (double) rhoRef, (double) mu);
  for(int i=0; i<8; ++i) std::printf("%s%.17g", i ? ", " : "", (double) monitoringData[i]);
  std::printf("], \"step_seconds\": [");
  for(size_t i=0; i<times.size(); ++i) std::printf("%s%.6g", i ? ", " : "", times[i]);
  std::printf("]}\n");
  return 0;
}
