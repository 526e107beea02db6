"""CPU FP64 restatement of the cubez step: TEST INFRASTRUCTURE ONLY (the
checker for the HIP path; never imported by the product)."""
