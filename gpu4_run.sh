mkdir -p gpurun_out
timeout -k 10 400 python -u -m pytest tests/test_gpu_parity.py -v --timeout 120 --timeout-method thread > gpurun_out/gpu4_tests.log 2>&1
rc=$?
echo "pytest rc=$rc"
grep -E "PASS|FAIL|Error|error" gpurun_out/gpu4_tests.log | tail -30
if [ $rc -ne 0 ] && [ $rc -ne 1 ]; then exit $rc; fi
echo "=== bench"
timeout -k 10 400 python -u bench.py --steps 10 --warmup 2 --plain-steps 3 --cpu-seconds 5 > gpurun_out/gpu4_bench.log 2>&1
rc2=$?
tail -6 gpurun_out/gpu4_bench.log
exit $rc2
