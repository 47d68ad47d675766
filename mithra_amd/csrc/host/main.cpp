// `mithra <job-file>`: the reference's command line (src/mithra.cpp:32-102) with the time
// march on the GPU.  Errors print their message and exit with status 1, as the reference does.
#include <sys/time.h>

#include <iostream>

#include "mithra_host.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::cout << "usage: mithra <job-file>" << std::endl;
    return 1;
  }
  timeval start, end;
  gettimeofday(&start, nullptr);
  mithra_job* job = nullptr;
  int rc = mithra_job_open(argv[1], &job);
  if (rc == 0) rc = mithra_job_initialize(job);
  long steps = rc == 0 ? mithra_job_run(job, -1, MITHRA_PRECISION_FP64, 0) : rc;
  if (rc != 0 || steps < 0) {
    std::cout << (job ? mithra_job_error(job) : "cannot open the job") << std::endl;
    mithra_job_close(job);
    return 1;
  }
  mithra_job_close(job);
  gettimeofday(&end, nullptr);
  double dt = (end.tv_usec - start.tv_usec) / 1.0e6 + (end.tv_sec - start.tv_sec);
  std::cout << "::: total simulation time [seconds] = " << dt << std::endl;
  return 0;
}
