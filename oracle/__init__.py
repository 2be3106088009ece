"""CPU restatement of DOMA's fixed-source sweep and source iteration.

TEST INFRASTRUCTURE ONLY (checker and CPU baseline); the product never imports it.
"""
