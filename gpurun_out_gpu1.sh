set -o pipefail
cd $GRAFT_REPO_ROOT
mkdir -p gpurun_out
echo "== gpu tests" 
timeout -k 10 420 python -u -m pytest tests/test_gpu_parity.py -x -v --timeout 300 --timeout-method thread > gpurun_out/gpu_tests.log 2>&1
rc=$?
tail -25 gpurun_out/gpu_tests.log
exit $rc
